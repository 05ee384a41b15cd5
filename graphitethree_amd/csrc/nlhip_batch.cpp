// Batched solves of many independent systems (texture-atlas charts): host
// side of k_batch.hip and its C ABI (include/nl_hip.h, batched part).
//
// The systems arrive as one block-diagonal CRS.  They are cut into tiles of
// kTileRows rows; a fixed grid of workgroups walks the active tiles; a
// captured hipGraph holds `check_every` iterations followed by a rebuild of
// the active lists, and the host only reads the number of active systems
// between graph launches.
#include <cstring>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>

#include "nlhip_internal.h"

namespace nlhip {

template <typename T>
struct BatchWork {
    DevBuf<T> x, b, r, h, d, q, part, sys;  // part: 4 x ntiles; sys: 7 x nsys
    DevBuf<int32_t> iters, done, stop, skip, cap;
    DevBuf<T> chk;                          // 2 x nsys: |b - A x|^2, |b|^2 (confirm)
    DevBuf<uint32_t> lists, counts;         // act_tiles | act_sys ; 2 counters
    PinnedBuf<uint32_t> host_counts;        // two slots
    GraphLadder graph;
    BatchArgs<T> args{};
    BatchArgs<T> graph_args{};
    int graph_max_iter = -1;
    bool graph_precond = false;
};

struct Batch {
    uint32_t nsys = 0, n = 0, ntiles = 0;
    uint64_t nnz = 0;
    DevBuf<uint32_t> row_off, tile_off, tile_sys, tile_row0, rowptr, colind;
    DevBuf<uint16_t> ci16;  // 16-bit column offsets (empty when they do not fit)
    DevBuf<uint32_t> segb;
    // 2x2 block form of the systems that have it (empty when none has)
    DevBuf<uint32_t> sys_pair0, brp, bcol, bsegb;
    bool all_blocks = false;  // every system has the block form
    DevBuf<uint16_t> bci16;
    DevBuf<double> bval;
    DevBuf<float> bval32;
    DevBuf<double> val, dinv, scale;  // scale: 2 x nsys (scale, unscale)
    DevBuf<int32_t> flags;            // per system, host-set (refinement undo)
    DevBuf<float> val32, dinv32;
    std::vector<uint32_t> h_row_off;
    uint32_t grid = 1;
    uint32_t spmv_passes = 0;  // 0: by active tiles, else 1, 2 or 4
    bool graphs = true;        // false: the same launches enqueued directly
    bool persistent = false;   // true: each run of iterations is one cooperative launch
    uint32_t last_fallbacks = 0;  // systems the last fallback solve finished in FP64
    Stream stream;
    std::mutex mu;
    BatchWork<double> w64;
    BatchWork<float> w32;
};

namespace {

constexpr uint32_t kSpmvR2PerCu = 16;
constexpr uint32_t kSpmvR4PerCu = 32;

// SpMV run length: with R passes per wavefront there are 4 n / R wavefront
// tasks for n active tiles; longer runs once that is well above the
// wavefronts a launch keeps resident (about 16 per CU).  A forced length
// (nlhipBatchSetSpmvPasses) applies from any number of tiles.
template <typename T>
void set_spmv_runs(const Batch& B, BatchArgs<T>& a) {
    const uint32_t cu = uint32_t(std::max(1, device().cu_count));
    constexpr uint32_t kNever = 0xffffffffu;
    const uint32_t p = B.spmv_passes;
    a.spmv_r2_tiles = p == 0 ? kSpmvR2PerCu * cu : (p >= 2 ? 0u : kNever);
    a.spmv_r4_tiles = p == 0 ? kSpmvR4PerCu * cu : (p == 4 ? 0u : kNever);
}

template <typename T>
void work_alloc(Batch& B, BatchWork<T>& w) {
    if (w.x.get()) return;
    const size_t n = B.n;
    w.x.alloc(n);
    w.b.alloc(n);
    w.r.alloc(n);
    w.h.alloc(n);
    w.d.alloc(n);
    w.q.alloc(n);
    w.part.alloc(4 * size_t(B.ntiles));
    w.sys.alloc(7 * size_t(B.nsys));
    w.iters.alloc(B.nsys);
    w.done.alloc(B.nsys);
    w.stop.alloc(B.nsys);
    w.skip.alloc(B.nsys);
    w.cap.alloc(B.nsys);
    w.chk.alloc(2 * size_t(B.nsys));
    w.lists.alloc(size_t(B.ntiles) + B.nsys);
    w.counts.alloc(2);
    w.host_counts.ensure(2);
    NLHIP_CHECK(fill<int32_t>(B.nsys, 0, w.skip.get(), B.stream.get()));
    BatchArgs<T>& a = w.args;
    a.nsys = B.nsys;
    a.ntiles = B.ntiles;
    a.row_off = B.row_off.get();
    a.tile_off = B.tile_off.get();
    a.tile_sys = B.tile_sys.get();
    a.tile_row0 = B.tile_row0.get();
    a.rp = B.rowptr.get();
    a.ci = B.colind.get();
    a.ci16 = B.ci16.get();
    a.segb = B.segb.get();
    a.sys_pair0 = B.sys_pair0.get();
    a.all_blocks = B.all_blocks ? 1u : 0u;
    set_spmv_runs(B, a);
    a.brp = B.brp.get();
    a.bcol = B.bcol.get();
    a.bci16 = B.bci16.get();
    a.bsegb = B.bsegb.get();
    if constexpr (std::is_same<T, double>::value) a.bval = B.bval.get();
    else a.bval = B.bval32.get();
    a.x = w.x.get();
    a.b = w.b.get();
    a.r = w.r.get();
    a.h = w.h.get();
    a.d = w.d.get();
    a.q = w.q.get();
    a.part_dq = w.part.get();
    a.part_rh = w.part.get() + B.ntiles;
    a.part_rr = w.part.get() + 2 * size_t(B.ntiles);
    a.part_bb = w.part.get() + 3 * size_t(B.ntiles);
    a.rho = w.sys.get();
    a.alpha = w.sys.get() + B.nsys;
    a.beta = w.sys.get() + 2 * size_t(B.nsys);
    a.err = w.sys.get() + 3 * size_t(B.nsys);
    a.thresh = w.sys.get() + 4 * size_t(B.nsys);
    a.bb = w.sys.get() + 5 * size_t(B.nsys);
    a.rho_prev = w.sys.get() + 6 * size_t(B.nsys);
    a.iters = w.iters.get();
    a.done = w.done.get();
    a.stop = w.stop.get();
    a.skip = nullptr;
    a.cap = nullptr;
    a.act_tiles = w.lists.get();
    a.act_sys = w.lists.get() + B.ntiles;
    a.n_act_tiles = w.counts.get();
    a.n_act_sys = w.counts.get() + 1;
}

// Runs the batched PCG (started, or re-opened with its lists rebuilt) until
// no system is active.  Graphs hold K iterations and a list rebuild;
// `budget` bounds the iterations any system can still run.
template <typename T>
void batch_loop(Batch& B, BatchWork<T>& w, bool precond, int max_iter, uint32_t K,
                uint64_t budget) {
    hipStream_t s = B.stream.get();
    // iterations per launch: 4, 4, 8, 16, ... up to K, as the graph ladder
    auto run_len = [&](uint64_t i) {
        uint32_t k = K < 4u ? K : 4u;
        for (uint64_t l = 1; l < i && k < K; ++l) k = 2 * k < K ? 2 * k : K;
        return k;
    };
    uint32_t pgrid = 0;
    if (B.persistent) {
        NLHIP_CHECK(batch_persistent_grid<T>(w.args, precond, &pgrid));
        pgrid = std::min(pgrid, std::max(B.grid, 1u));
    } else {
        if (w.graph_precond != precond || w.graph_max_iter != max_iter ||
            std::memcmp(&w.graph_args, &w.args, sizeof(BatchArgs<T>)) != 0)
            w.graph.reset();
        w.graph.ensure(K, s, [&](uint32_t k) {
            for (uint32_t i = 0; i < k; ++i)
                NLHIP_CHECK(batch_iteration<T>(w.args, precond, max_iter, B.grid, s));
            NLHIP_CHECK(batch_compact<T>(w.args, B.grid, s));
        });
        w.graph_precond = precond;
        w.graph_max_iter = max_iter;
        w.graph_args = w.args;
    }
    Event ev[2];
    uint32_t* host = w.host_counts.get();
    const uint64_t limit = budget + 3ull * K + 8;
    uint64_t launched = 0, covered = 0;
    for (;;) {
        const uint32_t its = run_len(launched);
        if (B.persistent) {
            NLHIP_CHECK(batch_persistent<T>(w.args, precond, max_iter, its, pgrid, s));
        } else if (B.graphs) {
            w.graph.launch(launched, s);
        } else {
            for (uint32_t i = 0; i < its; ++i)
                NLHIP_CHECK(batch_iteration<T>(w.args, precond, max_iter, B.grid, s));
            NLHIP_CHECK(batch_compact<T>(w.args, B.grid, s));
        }
        covered += its;
        const int slot = int(launched & 1);
        NLHIP_CHECK(hipMemcpyAsync(&host[slot], w.counts.get() + 1, sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s));
        ev[slot].record(s);
        ++launched;
        if (launched >= 2) {
            const int prev = int((launched - 2) & 1);
            ev[prev].sync();
            if (host[prev] == 0) break;
        }
        if (covered > limit) {
            B.stream.sync();
            if (host[slot] == 0) break;
            throw Error(NLHIP_ERR_HIP, "batched CG did not reach its stopping test");
        }
    }
    B.stream.sync();
}

template <typename T>
void run_batch(Batch& B, BatchWork<T>& w, bool precond, T eps, int max_iter, uint32_t K) {
    NLHIP_CHECK(batch_start<T>(w.args, precond, eps, max_iter, B.grid, B.stream.get()));
    batch_loop<T>(B, w, precond, max_iter, K, uint64_t(std::max(max_iter, 0)));
}

uint32_t batch_check_every(const Batch& B) {
    const double bytes = 12.0 * double(B.nnz) + 108.0 * double(B.n);
    const double iter_s = std::max(bytes / 4.0e12, 3.0e-5);
    return uint32_t(std::clamp(4.0e-3 / iter_s, 4.0, 64.0));
}

void ensure_fp32(Batch& B) {
    if (B.val32.get()) return;
    B.val32.alloc(B.nnz);
    B.dinv32.alloc(B.n);
    NLHIP_CHECK(cast_f64_f32(B.nnz, B.val.get(), B.val32.get(), B.stream.get()));
    NLHIP_CHECK(jacobi_dinv<float>(B.n, B.rowptr.get(), B.colind.get(), B.val.get(),
                                   B.dinv32.get(), B.stream.get()));
    if (B.bval.get()) {
        B.bval32.alloc(B.bval.size());
        NLHIP_CHECK(block_vals_f32(B.bval.size() / 4, B.bval.get(), B.bval32.get(),
                                   B.stream.get()));
    }
}

void fill_stats(Batch& B, const std::vector<int32_t>& its, const std::vector<double>& rel,
                const std::vector<double>& cgrel, const std::vector<uint8_t>& conv,
                const std::vector<uint32_t>* outer, const std::vector<uint32_t>& extra,
                nlhip_solve_stats* per, nlhip_solve_stats* total, double dev_s, double wall_s) {
    uint64_t it_sum = 0;
    uint32_t it_max = 0, nconv = 0, extra_max = 0;
    double worst = 0.0;
    for (uint32_t s = 0; s < B.nsys; ++s) {
        it_sum += uint64_t(its[s]);
        it_max = std::max(it_max, uint32_t(its[s]));
        nconv += conv[s];
        worst = std::max(worst, rel[s]);
        const uint32_t ex = extra.empty() ? 0u : extra[s];
        extra_max = std::max(extra_max, ex);
        if (per) {
            nlhip_solve_stats& st = per[s];
            std::memset(&st, 0, sizeof(st));
            st.iterations = uint32_t(its[s]);
            st.outer_iterations = outer ? (*outer)[s] : 0;
            st.converged = conv[s];
            st.rel_residual = rel[s];
            st.cg_rel_residual = cgrel[s];
            st.device_seconds = dev_s;
            st.wall_seconds = wall_s;
            st.extra_iterations = ex;
        }
    }
    if (total) {
        std::memset(total, 0, sizeof(*total));
        total->iterations = it_max;
        total->outer_iterations = uint32_t(std::min<uint64_t>(it_sum, 0xFFFFFFFFu));
        total->converged = (nconv == B.nsys) ? 1 : 0;
        total->rel_residual = worst;
        total->device_seconds = dev_s;
        total->wall_seconds = wall_s;
        total->extra_iterations = extra_max;
    }
}

// FP64 |b - A x| / |b| for every system.
void true_residuals(Batch& B, std::vector<double>& rel, std::vector<double>* rr_out = nullptr,
                    std::vector<double>* bb_out = nullptr) {
    work_alloc(B, B.w64);
    BatchArgs<double> a = B.w64.args;
    a.val = B.val.get();
    NLHIP_CHECK(batch_residual64(a, B.grid, B.stream.get()));
    std::vector<double> rr(B.nsys), bb(B.nsys);
    NLHIP_CHECK(hipMemcpyAsync(rr.data(), a.err, B.nsys * 8, hipMemcpyDeviceToHost,
                               B.stream.get()));
    NLHIP_CHECK(hipMemcpyAsync(bb.data(), a.bb, B.nsys * 8, hipMemcpyDeviceToHost,
                               B.stream.get()));
    B.stream.sync();
    rel.resize(B.nsys);
    for (uint32_t s = 0; s < B.nsys; ++s)
        rel[s] = bb[s] > 0 ? std::sqrt(rr[s] / bb[s]) : std::sqrt(rr[s]);
    if (rr_out) *rr_out = rr;
    if (bb_out) *bb_out = bb;
}

// |b - A x|^2 and |b|^2 per system without touching what a stopped FP64 CG
// needs to continue (r, the r.h / r.r tile partials, the per-system
// scalars): the residual goes to q and the d.q partials, the norms to chk.
void confirm_norms(Batch& B, std::vector<double>& rr, std::vector<double>& bb) {
    BatchWork<double>& w = B.w64;
    BatchArgs<double> a = w.args;
    a.val = B.val.get();
    a.r = w.q.get();
    a.part_rr = a.part_dq;
    a.err = w.chk.get();
    a.bb = w.chk.get() + B.nsys;
    NLHIP_CHECK(batch_residual64(a, B.grid, B.stream.get()));
    rr.resize(B.nsys);
    bb.resize(B.nsys);
    NLHIP_CHECK(hipMemcpyAsync(rr.data(), a.err, B.nsys * 8, hipMemcpyDeviceToHost,
                               B.stream.get()));
    NLHIP_CHECK(hipMemcpyAsync(bb.data(), a.bb, B.nsys * 8, hipMemcpyDeviceToHost,
                               B.stream.get()));
    B.stream.sync();
}

constexpr int kMaxConfirmRounds = 8;

// FP64 batched PCG from the current x of every system that skip (host, may
// be null) does not exclude, each system stopping at its own cap (host, may
// be null: p.max_iter).  With p.confirm, a system whose recursive residual
// met eps but whose true residual |b - A x| did not continues the same
// recurrence (threshold scaled down by the ratio measured, alpha = 0, its
// loop re-opened) until the true residual meets eps, its cap is reached or
// the true residual stops decreasing; the other systems stay stopped, so
// each system's iterates are those of its solve alone.  Results for the
// systems run: its, cgrel (recursive), conv, extra (iterations after the
// recursive test first held).
void batch_fp64(Batch& B, const nlhip_solve_params& p, const std::vector<int32_t>* skip,
                const std::vector<int32_t>* cap, std::vector<int32_t>& its,
                std::vector<double>& cgrel, std::vector<uint8_t>& conv,
                std::vector<uint32_t>& extra) {
    BatchWork<double>& w = B.w64;
    hipStream_t s = B.stream.get();
    const uint32_t ns = B.nsys;
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    w.args.val = B.val.get();
    w.args.dinv = precond ? B.dinv.get() : nullptr;
    w.args.skip = nullptr;
    w.args.cap = nullptr;
    if (skip) {
        NLHIP_CHECK(hipMemcpyAsync(w.skip.get(), skip->data(), ns * 4, hipMemcpyHostToDevice, s));
        w.args.skip = w.skip.get();
    }
    if (cap) {
        NLHIP_CHECK(hipMemcpyAsync(w.cap.get(), cap->data(), ns * 4, hipMemcpyHostToDevice, s));
        w.args.cap = w.cap.get();
    }
    const uint32_t K = p.check_every ? p.check_every : batch_check_every(B);
    run_batch<double>(B, w, precond, p.eps, int(p.max_iter), K);
    std::vector<double> err(ns), thr(ns), bb(ns);
    its.resize(ns);
    auto read = [&] {
        NLHIP_CHECK(hipMemcpyAsync(its.data(), w.iters.get(), ns * 4, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipMemcpyAsync(err.data(), w.args.err, ns * 8, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipMemcpyAsync(thr.data(), w.args.thresh, ns * 8, hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(hipMemcpyAsync(bb.data(), w.args.bb, ns * 8, hipMemcpyDeviceToHost, s));
        B.stream.sync();
    };
    read();
    auto ran = [&](uint32_t k) { return !(skip && (*skip)[k]); };
    auto cap_of = [&](uint32_t k) {
        return cap ? std::min<int64_t>((*cap)[k], int64_t(p.max_iter)) : int64_t(p.max_iter);
    };
    std::vector<uint32_t> first(ns, ~0u);
    for (uint32_t k = 0; k < ns; ++k) {
        if (!ran(k)) continue;
        if (!std::isfinite(err[k]) || !std::isfinite(bb[k]))
            throw Error(NLHIP_ERR_NONFINITE, "system " + std::to_string(k) +
                                                 ": right-hand side or iterate is NaN or Inf");
        if (!(err[k] > thr[k])) first[k] = uint32_t(its[k]);
    }
    std::vector<double> rr, tb;
    if (p.confirm) {
        std::vector<double> last(ns, -1.0), alpha(ns);
        std::vector<int32_t> stop(ns), done(ns);
        for (int round = 0;; ++round) {
            confirm_norms(B, rr, tb);
            std::vector<uint32_t> again;
            for (uint32_t k = 0; k < ns; ++k) {
                if (!ran(k) || err[k] > thr[k]) continue;
                if (first[k] == ~0u) first[k] = uint32_t(its[k]);
                const double target = (p.eps * p.eps) * tb[k];
                if (rr[k] > target && its[k] < cap_of(k) && err[k] > 0.0 &&
                    (last[k] < 0.0 || rr[k] <= 0.99 * last[k]) && round < kMaxConfirmRounds)
                    again.push_back(k);
            }
            if (again.empty()) break;
            NLHIP_CHECK(hipMemcpyAsync(alpha.data(), w.args.alpha, ns * 8, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(hipMemcpyAsync(stop.data(), w.stop.get(), ns * 4, hipMemcpyDeviceToHost, s));
            NLHIP_CHECK(hipMemcpyAsync(done.data(), w.done.get(), ns * 4, hipMemcpyDeviceToHost, s));
            B.stream.sync();
            for (uint32_t k : again) {
                const double target = (p.eps * p.eps) * tb[k];
                thr[k] = err[k] * (target / rr[k]) * 0.5;
                alpha[k] = 0.0;  // the launch that stopped applied x -= alpha d
                stop[k] = 0;
                done[k] = 0;
                last[k] = rr[k];
            }
            NLHIP_CHECK(hipMemcpyAsync(w.args.thresh, thr.data(), ns * 8, hipMemcpyHostToDevice, s));
            NLHIP_CHECK(hipMemcpyAsync(w.args.alpha, alpha.data(), ns * 8, hipMemcpyHostToDevice, s));
            NLHIP_CHECK(hipMemcpyAsync(w.stop.get(), stop.data(), ns * 4, hipMemcpyHostToDevice, s));
            NLHIP_CHECK(hipMemcpyAsync(w.done.get(), done.data(), ns * 4, hipMemcpyHostToDevice, s));
            NLHIP_CHECK(batch_compact<double>(w.args, B.grid, s));
            batch_loop<double>(B, w, precond, int(p.max_iter), K, p.max_iter);
            read();
        }
    }
    cgrel.resize(ns);
    conv.resize(ns);
    extra.resize(ns);
    for (uint32_t k = 0; k < ns; ++k) {
        if (!ran(k)) continue;
        cgrel[k] = bb[k] > 0 ? std::sqrt(err[k] / bb[k]) : std::sqrt(err[k]);
        conv[k] = p.confirm ? (first[k] != ~0u && !(rr[k] > (p.eps * p.eps) * tb[k]) ? 1 : 0)
                            : (!(err[k] > thr[k]) ? 1 : 0);
        extra[k] = first[k] == ~0u ? 0u : uint32_t(its[k]) - first[k];
    }
    w.args.skip = nullptr;
    w.args.cap = nullptr;
}

void batch_solve_fp64(Batch& B, const nlhip_solve_params& p, std::vector<int32_t>& its,
                      std::vector<double>& cgrel, std::vector<uint8_t>& conv,
                      std::vector<uint32_t>& extra) {
    batch_fp64(B, p, nullptr, nullptr, its, cgrel, conv, extra);
}

void batch_solve_mixed(Batch& B, const nlhip_solve_params& p, std::vector<int32_t>& its,
                       std::vector<double>& cgrel, std::vector<uint8_t>& conv,
                       std::vector<uint32_t>& outer, std::vector<uint8_t>* fell_back = nullptr) {
    ensure_fp32(B);
    work_alloc(B, B.w32);
    BatchWork<double>& w64 = B.w64;
    BatchWork<float>& w32 = B.w32;
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    w32.args.val = B.val32.get();
    w32.args.dinv = precond ? B.dinv32.get() : nullptr;
    w32.args.skip = w32.skip.get();
    w32.args.cap = w32.cap.get();
    hipStream_t s = B.stream.get();
    B.scale.ensure(2 * size_t(B.nsys));
    const uint32_t K = p.check_every ? p.check_every : batch_check_every(B);
    const uint32_t inner_cap = p.inner_max_iter ? p.inner_max_iter : p.max_iter;
    its.assign(B.nsys, 0);
    outer.assign(B.nsys, 0);
    conv.assign(B.nsys, 0);
    cgrel.assign(B.nsys, 0.0);
    std::vector<int32_t> skip(B.nsys), inner_its(B.nsys), cap(B.nsys);
    std::vector<int32_t> refined(B.nsys, 0), save(B.nsys), back(B.nsys), stalled(B.nsys, 0);
    std::vector<uint32_t> no_gain(B.nsys, 0);
    std::vector<double> sc(2 * size_t(B.nsys)), rr_best(B.nsys, 0.0);
    // each system's iterate with the lowest FP64 residual so far (the FP64
    // CG's d is free here); a system's refinement ends back at it
    // (kRefineNoGain / kRefineDiverged, as for one system), and with the
    // fallback it continues in FP64 from there
    double* const x_best = w64.d.get();
    B.flags.ensure(B.nsys);
    bool first = true, finish = false;
    for (;;) {
        std::vector<double> rel, rr, bb;
        true_residuals(B, rel, &rr, &bb);  // leaves r64 = b - A x in w64.r
        if (first)
            for (uint32_t k = 0; k < B.nsys; ++k)
                if (!std::isfinite(rr[k]) || !std::isfinite(bb[k]))
                    throw Error(NLHIP_ERR_NONFINITE,
                                "system " + std::to_string(k) +
                                    ": right-hand side or iterate is NaN or Inf");
        uint32_t active = 0;
        bool any_save = false, any_back = false;
        for (uint32_t k = 0; k < B.nsys; ++k) {
            save[k] = back[k] = 0;
            if (first || (refined[k] && rr[k] < rr_best[k])) {
                save[k] = 1;
                rr_best[k] = rr[k];
                no_gain[k] = 0;
            } else if (refined[k] && (!std::isfinite(rr[k]) ||
                                      rr[k] > kRefineDiverged * rr_best[k] ||
                                      ++no_gain[k] >= kRefineNoGain)) {
                stalled[k] = 1;
                if (fell_back && uint32_t(its[k]) < p.max_iter) (*fell_back)[k] = 1;
            }
            const bool budget = uint32_t(its[k]) >= p.max_iter || outer[k] >= p.max_outer;
            const bool gone = fell_back && (*fell_back)[k];
            const bool done = !(rr[k] > (p.eps * p.eps) * bb[k]) || budget || gone ||
                              stalled[k] || finish;
            if (done && !(rr[k] <= rr_best[k])) {  // back to the best iterate
                back[k] = 1;
                rr[k] = rr_best[k];
                rel[k] = bb[k] > 0 ? std::sqrt(rr[k] / bb[k]) : std::sqrt(rr[k]);
            }
            any_save |= save[k] != 0;
            any_back |= back[k] != 0;
            conv[k] = !(rr[k] > (p.eps * p.eps) * bb[k]) ? 1 : 0;
            cgrel[k] = rel[k];
            skip[k] = done ? 1 : 0;
            refined[k] = done ? 0 : 1;
            // each inner solve is capped by what is left of the system's budget
            cap[k] = int32_t(std::min<uint32_t>(inner_cap, p.max_iter - std::min<uint32_t>(
                                                                      p.max_iter, its[k])));
            const double rn = std::sqrt(rr[k]);
            sc[k] = skip[k] ? 0.0 : 1.0 / rn;
            sc[B.nsys + k] = skip[k] ? 0.0 : rn;
            active += skip[k] ? 0 : 1;
        }
        first = false;
        if (any_save) {
            NLHIP_CHECK(hipMemcpyAsync(B.flags.get(), save.data(), B.nsys * 4,
                                       hipMemcpyHostToDevice, s));
            NLHIP_CHECK(batch_copy_systems(w64.args, x_best, w64.x.get(), B.flags.get(), B.grid, s));
        }
        if (any_back) {
            NLHIP_CHECK(hipMemcpyAsync(B.flags.get(), back.data(), B.nsys * 4,
                                       hipMemcpyHostToDevice, s));
            NLHIP_CHECK(batch_copy_systems(w64.args, w64.x.get(), x_best, B.flags.get(), B.grid, s));
        }
        if (active == 0) break;
        NLHIP_CHECK(hipMemcpyAsync(w32.skip.get(), skip.data(), B.nsys * 4, hipMemcpyHostToDevice, s));
        NLHIP_CHECK(hipMemcpyAsync(w32.cap.get(), cap.data(), B.nsys * 4, hipMemcpyHostToDevice, s));
        NLHIP_CHECK(hipMemcpyAsync(B.scale.get(), sc.data(), sc.size() * 8, hipMemcpyHostToDevice, s));
        NLHIP_CHECK(batch_scale_cast(w64.args, w32.args, B.scale.get(), B.grid, s));
        run_batch<float>(B, w32, precond, float(p.inner_eps), int(inner_cap), K);
        NLHIP_CHECK(hipMemcpyAsync(inner_its.data(), w32.iters.get(), B.nsys * 4,
                                   hipMemcpyDeviceToHost, s));
        NLHIP_CHECK(batch_ir_update(w64.args, w32.args, B.scale.get() + B.nsys, B.grid, s));
        B.stream.sync();
        bool progress = false;
        for (uint32_t k = 0; k < B.nsys; ++k) {
            if (skip[k]) continue;
            its[k] += inner_its[k];
            outer[k] += 1;
            progress |= inner_its[k] > 0;
            // an inner solve that used its whole cap without reaching
            // inner_eps while budget is left: this system continues in FP64
            if (fell_back && inner_its[k] >= cap[k] && uint32_t(its[k]) < p.max_iter)
                (*fell_back)[k] = 1;
        }
        finish = !progress;  // the next pass settles every system and ends
    }
    w32.args.cap = nullptr;
}

// Mixed precision with a fallback (the policy of LAPACK's DSGESV: when the
// FP32 inner solver does not converge, finish in FP64): systems whose FP32
// inner solve reaches its cap without reaching inner_eps stop refining and
// are solved by the FP64 batched CG from their current iterate, within what
// is left of their max_iter; the others refine to eps as in
// batch_solve_mixed.
void batch_solve_mixed_fallback(Batch& B, const nlhip_solve_params& p, std::vector<int32_t>& its,
                                std::vector<double>& cgrel, std::vector<uint8_t>& conv,
                                std::vector<uint32_t>& outer, std::vector<uint32_t>& extra,
                                uint32_t& nfallback) {
    if (p.inner_max_iter == 0)
        throw Error(NLHIP_ERR_INVALID, "the fallback needs inner_max_iter > 0");
    std::vector<uint8_t> fb(B.nsys, 0);
    // systems that fall back are dropped from the refinement at once: the
    // FP32 loop skips them from the next outer step on
    batch_solve_mixed(B, p, its, cgrel, conv, outer, &fb);
    extra.assign(B.nsys, 0);
    nfallback = 0;
    std::vector<int32_t> skip(B.nsys), cap(B.nsys);
    for (uint32_t k = 0; k < B.nsys; ++k) {
        const bool go = fb[k] && !conv[k];
        skip[k] = go ? 0 : 1;
        cap[k] = int32_t(p.max_iter - std::min<uint32_t>(p.max_iter, its[k]));
        nfallback += go ? 1u : 0u;
    }
    if (nfallback == 0) return;
    std::vector<int32_t> its64;
    std::vector<double> cgrel64;
    std::vector<uint8_t> conv64;
    std::vector<uint32_t> extra64;
    batch_fp64(B, p, &skip, &cap, its64, cgrel64, conv64, extra64);
    for (uint32_t k = 0; k < B.nsys; ++k) {
        if (skip[k]) continue;
        its[k] += its64[k];
        cgrel[k] = cgrel64[k];
        conv[k] = conv64[k];
        extra[k] = extra64[k];
    }
}

// 2x2 block form per system (see BlockStreamEngine): system s takes it when
// it has an even number of rows and every row pair (row_off[s] + 2p, +1) of
// it has the same even-length column list made of pairs (c, c + 1) with
// c - row_off[s] even.  Whether a system does depends only on that system
// (not on its offset in the batch), so its tiles, and its result, are the
// same in any batch.  Built on the host (the arrays are host arrays here);
// block columns as 16-bit offsets when the whole block stream fits them,
// 32-bit otherwise (same sums either way).
void build_batch_blocks(Batch& B, const uint32_t* row_off, const uint32_t* rowptr,
                        const uint32_t* colind, const double* val, std::vector<int32_t>& blk) {
    const uint32_t nsys = B.nsys;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < int64_t(nsys); ++s) {
        const uint32_t lo = row_off[s], hi = row_off[s + 1];
        bool ok = ((hi - lo) % 2 == 0);
        for (uint32_t i = lo; ok && i < hi; i += 2) {
            const uint32_t b0 = rowptr[i], b1 = rowptr[i + 1], e1 = rowptr[i + 2];
            const uint32_t len = b1 - b0;
            ok = (e1 - b1 == len) && (len % 2 == 0);
            for (uint32_t q = 0; ok && q < len; q += 2) {
                const uint32_t c = colind[b0 + q];
                ok = (c - lo) % 2 == 0 && colind[b0 + q + 1] == c + 1 && colind[b1 + q] == c &&
                     colind[b1 + q + 1] == c + 1;
            }
        }
        blk[size_t(s)] = ok ? 1 : 0;
    }
    std::vector<uint32_t> pair0(nsys, kNoBlocks);
    uint32_t np = 0;
    for (uint32_t s = 0; s < nsys; ++s)
        if (blk[s]) {
            pair0[s] = np;
            np += (row_off[s + 1] - row_off[s]) / 2;
        }
    if (np == 0) return;
    B.all_blocks = std::all_of(blk.begin(), blk.begin() + nsys, [](int32_t v) { return v != 0; });
    std::vector<uint32_t> h_brp(size_t(np) + 1, 0);
    for (uint32_t s = 0; s < nsys; ++s)
        if (blk[s])
            for (uint32_t i = row_off[s], B = pair0[s]; i < row_off[s + 1]; i += 2, ++B)
                h_brp[B + 1] = (rowptr[i + 1] - rowptr[i]) / 2;
    for (uint32_t I = 0; I < np; ++I) h_brp[I + 1] += h_brp[I];
    const uint64_t nblk = h_brp[np];
    std::vector<uint32_t> h_bcol(nblk);
    std::vector<double> h_bval(block_vals_size(nblk), 0.0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < int64_t(nsys); ++s) {
        if (!blk[size_t(s)]) continue;
        const uint32_t lo = row_off[s], p0 = pair0[size_t(s)];
        for (uint32_t i = lo, Bk = p0; i < row_off[s + 1]; i += 2, ++Bk) {
            const uint32_t b0 = rowptr[i], b1 = rowptr[i + 1], len = b1 - b0;
            uint64_t k = h_brp[Bk];
            for (uint32_t q = 0; q < len; q += 2, ++k) {
                h_bcol[k] = p0 + (colind[b0 + q] - lo) / 2;
                h_bval[block_upper<double>(k)] = val[b0 + q];
                h_bval[block_upper<double>(k) + 1] = val[b0 + q + 1];
                h_bval[block_lower<double>(k)] = val[b1 + q];
                h_bval[block_lower<double>(k) + 1] = val[b1 + q + 1];
            }
        }
    }
    hipStream_t st = B.stream.get();
    B.sys_pair0.alloc(nsys);
    B.brp.alloc(size_t(np) + 1);
    B.bcol.alloc(nblk);
    B.bval.alloc(h_bval.size());
    copy_in(B.sys_pair0.get(), pair0.data(), size_t(nsys) * 4, st);
    copy_in(B.brp.get(), h_brp.data(), h_brp.size() * 4, st);
    copy_in(B.bcol.get(), h_bcol.data(), nblk * 4, st);
    copy_in(B.bval.get(), h_bval.data(), h_bval.size() * 8, st);
    if (build_col_offsets(nblk, B.bcol.get(), st, B.bci16, B.bsegb)) B.bcol.release();
    NLHIP_CHECK(hipStreamSynchronize(st));  // host staging arrays go out of scope
}

void batch_finish(Batch& B, const std::vector<int32_t>& blk, uint32_t flags);
bool build_batch_blocks_device(Batch& B, std::vector<int32_t>& blk);

Batch* batch_create(uint32_t nsys, const uint32_t* row_off, const uint32_t* rowptr,
                    const uint32_t* colind, const double* val, uint32_t flags) {
    if (nsys == 0 || !row_off || !rowptr) throw Error(NLHIP_ERR_INVALID, "empty batch");
    if (row_off[0] != 0) throw Error(NLHIP_ERR_INVALID, "row_off[0] != 0");
    const uint32_t n = row_off[nsys];
    if (n == 0) throw Error(NLHIP_ERR_INVALID, "batch without rows");
    for (uint32_t s = 0; s < nsys; ++s)
        if (row_off[s + 1] <= row_off[s])
            throw Error(NLHIP_ERR_INVALID, "system " + std::to_string(s) + " has no rows");
    if (rowptr[0] != 0) throw Error(NLHIP_ERR_INVALID, "rowptr[0] != 0");
    for (uint32_t i = 0; i < n; ++i)
        if (rowptr[i + 1] < rowptr[i]) throw Error(NLHIP_ERR_INVALID, "rowptr is not monotone");
    const uint64_t nnz = rowptr[n];
    // every entry must stay inside its own system's diagonal block
    int bad_col = 0, bad_val = 0;
#pragma omp parallel for reduction(| : bad_col, bad_val) schedule(dynamic, 16)
    for (int64_t s = 0; s < int64_t(nsys); ++s) {
        const uint32_t lo = row_off[s], hi = row_off[s + 1];
        for (uint64_t j = rowptr[lo]; j < rowptr[hi]; ++j) {
            bad_col |= (colind[j] < lo || colind[j] >= hi);
            bad_val |= !std::isfinite(val[j]);
        }
    }
    if (bad_col) throw Error(NLHIP_ERR_INVALID, "an entry couples two systems (not block diagonal)");
    if (bad_val) throw Error(NLHIP_ERR_NONFINITE, "matrix value is NaN or Inf");
    device();
    auto B = std::make_unique<Batch>();
    B->nsys = nsys;
    B->n = n;
    B->nnz = nnz;
    B->h_row_off.assign(row_off, row_off + nsys + 1);
    hipStream_t st = B->stream.get();
    B->row_off.alloc(size_t(nsys) + 1);
    B->rowptr.alloc(size_t(n) + 1);
    B->colind.alloc(nnz);
    B->val.alloc(nnz);
    B->dinv.alloc(n);
    copy_in(B->row_off.get(), row_off, (size_t(nsys) + 1) * 4, st);
    copy_in(B->rowptr.get(), rowptr, (size_t(n) + 1) * 4, st);
    copy_in(B->colind.get(), colind, nnz * 4, st);
    copy_in(B->val.get(), val, nnz * 8, st);
    NLHIP_CHECK(jacobi_dinv<double>(n, B->rowptr.get(), B->colind.get(), B->val.get(),
                                    B->dinv.get(), st));
    std::vector<int32_t> blk(nsys, 0);
    if (!(flags & NLHIP_SYSTEM_NO_OFFSETS)) {
        build_col_offsets(nnz, B->colind.get(), st, B->ci16, B->segb);
        // every system in block form: checked and packed on the device for
        // the whole matrix at once; otherwise system by system on the host
        if (!(flags & NLHIP_SYSTEM_NO_BLOCKS) && !build_batch_blocks_device(*B, blk))
            build_batch_blocks(*B, row_off, rowptr, colind, val, blk);
    }
    batch_finish(*B, blk, flags);
    return B.release();
}

// The 2x2 block form of a resident batch whose systems all start at even
// rows, checked and packed on the device for the whole matrix at once (the
// single-system kernels): it holds when every global row pair (2I, 2I+1)
// has the pair structure, and then each system has it with its pairs at
// global pairs, sys_pair0[s] = row_off[s] / 2.  False (no block form) when
// a pair fails or a system starts at an odd row.
bool build_batch_blocks_device(Batch& B, std::vector<int32_t>& blk) {
    for (uint32_t s = 0; s <= B.nsys; ++s)
        if (B.h_row_off[s] & 1u) return false;
    hipStream_t st = B.stream.get();
    const uint32_t nb = B.n / 2;
    {
        DevBuf<uint32_t> bad(1);
        NLHIP_CHECK(hipMemsetAsync(bad.get(), 0, 4, st));
        NLHIP_CHECK(block_check(nb, B.rowptr.get(), B.colind.get(), bad.get(), st));
        uint32_t h = 0;
        NLHIP_CHECK(hipMemcpyAsync(&h, bad.get(), 4, hipMemcpyDeviceToHost, st));
        NLHIP_CHECK(hipStreamSynchronize(st));
        if (h) return false;
    }
    const uint64_t nblk = B.nnz / 4;
    std::vector<uint32_t> pair0(B.nsys);
    for (uint32_t s = 0; s < B.nsys; ++s) pair0[s] = B.h_row_off[s] / 2;
    B.sys_pair0.alloc(B.nsys);
    B.brp.alloc(size_t(nb) + 1);
    B.bcol.alloc(nblk);
    B.bval.alloc(block_vals_size(nblk));
    copy_in(B.sys_pair0.get(), pair0.data(), size_t(B.nsys) * 4, st);
    NLHIP_CHECK(block_pack<double>(nb, B.rowptr.get(), B.colind.get(), B.val.get(), B.brp.get(),
                                   B.bcol.get(), B.bval.get(), st));
    if (build_col_offsets(nblk, B.bcol.get(), st, B.bci16, B.bsegb)) B.bcol.release();
    NLHIP_CHECK(hipStreamSynchronize(st));
    B.all_blocks = true;
    std::fill(blk.begin(), blk.end(), 1);
    return true;
}

// Takes over a device CRS built by this library (systems [h_row_off[s],
// h_row_off[s+1]), entries inside their own diagonal blocks, sorted, finite)
// and finishes the batch as batch_create does.
Batch* batch_adopt(uint32_t nsys, const uint32_t* h_row_off, uint64_t nnz,
                   DevBuf<uint32_t>&& rowptr, DevBuf<uint32_t>&& colind, DevBuf<double>&& val,
                   uint32_t flags) {
    if (nsys == 0) throw Error(NLHIP_ERR_INVALID, "empty batch");
    for (uint32_t s = 0; s < nsys; ++s)
        if (h_row_off[s + 1] <= h_row_off[s])
            throw Error(NLHIP_ERR_INVALID, "system " + std::to_string(s) + " has no rows");
    auto B = std::make_unique<Batch>();
    B->nsys = nsys;
    B->n = h_row_off[nsys];
    B->nnz = nnz;
    B->h_row_off.assign(h_row_off, h_row_off + nsys + 1);
    hipStream_t st = B->stream.get();
    B->row_off.alloc(size_t(nsys) + 1);
    copy_in(B->row_off.get(), h_row_off, (size_t(nsys) + 1) * 4, st);
    B->rowptr = std::move(rowptr);
    B->colind = std::move(colind);
    B->val = std::move(val);
    B->dinv.alloc(B->n);
    NLHIP_CHECK(jacobi_dinv<double>(B->n, B->rowptr.get(), B->colind.get(), B->val.get(),
                                    B->dinv.get(), st));
    std::vector<int32_t> blk(nsys, 0);
    if (!(flags & NLHIP_SYSTEM_NO_OFFSETS)) {
        build_col_offsets(nnz, B->colind.get(), st, B->ci16, B->segb);
        if (!(flags & NLHIP_SYSTEM_NO_BLOCKS)) build_batch_blocks_device(*B, blk);
    }
    batch_finish(*B, blk, flags);
    return B.release();
}

// Tiles, launch grid, FP32 copy and FP64 work buffers of a batch whose
// matrix (and block form, per blk) is resident.
void batch_finish(Batch& Bref, const std::vector<int32_t>& blk, uint32_t flags) {
    Batch* B = &Bref;
    const uint32_t nsys = B->nsys;
    const uint32_t* row_off = B->h_row_off.data();
    hipStream_t st = B->stream.get();
    std::vector<uint32_t> tile_off(size_t(nsys) + 1, 0), tile_sys, tile_row0;
    for (uint32_t s = 0; s < nsys; ++s) {
        const uint32_t rows = row_off[s + 1] - row_off[s];
        const uint32_t TR = blk[s] ? kBlockTileRows : kTileRows;
        const uint32_t nt = (rows + TR - 1) / TR;
        tile_off[s + 1] = tile_off[s] + nt;
        for (uint32_t t = 0; t < nt; ++t) {
            tile_sys.push_back(s);
            tile_row0.push_back(row_off[s] + t * TR);
        }
    }
    B->ntiles = tile_off[nsys];
    B->tile_off.alloc(size_t(nsys) + 1);
    B->tile_sys.alloc(B->ntiles);
    B->tile_row0.alloc(B->ntiles);
    copy_in(B->tile_off.get(), tile_off.data(), tile_off.size() * 4, st);
    copy_in(B->tile_sys.get(), tile_sys.data(), tile_sys.size() * 4, st);
    copy_in(B->tile_row0.get(), tile_row0.data(), tile_row0.size() * 4, st);
    // 64 workgroups per CU, each taking a contiguous share of the active
    // tiles: several times what fits on a CU at once, so the dispatcher
    // balances shares of unequal cost; fewer when the batch has fewer tiles
    // (a small batch, e.g. one served request, launches no idle workgroups)
    B->grid = uint32_t(std::max<uint64_t>(
        1, std::min<uint64_t>(uint64_t(std::max(4, device().cu_count * 64)), B->ntiles)));
    if (flags & NLHIP_SYSTEM_FP32) ensure_fp32(*B);
    work_alloc(*B, B->w64);
    B->stream.sync();
}

}  // namespace

}  // namespace nlhip

using namespace nlhip;

static Batch* bat(nlhip_batch h) {
    if (!h) throw Error(NLHIP_ERR_INVALID, "null batch handle");
    return reinterpret_cast<Batch*>(h);
}

extern "C" {

NLAPI int NLAPIENTRY nlhipBatchCreate(uint32_t nsys, const uint32_t* row_off,
                                      const uint32_t* rowptr, const uint32_t* colind,
                                      const double* val, uint32_t flags, nlhip_batch* out) {
    return guarded("nlhipBatchCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        *out = reinterpret_cast<nlhip_batch>(batch_create(nsys, row_off, rowptr, colind, val, flags));
    });
}

NLAPI int NLAPIENTRY nlhipBatchDestroy(nlhip_batch b) {
    return guarded("nlhipBatchDestroy", [&] {
        if (!b) return;
        Batch* B = bat(b);
        { std::lock_guard<std::mutex> lock(B->mu); B->stream.sync(); }
        delete B;
    });
}

NLAPI int NLAPIENTRY nlhipBatchInfo(nlhip_batch b, uint32_t* nsys, uint32_t* n, uint64_t* nnz,
                                    uint32_t* tiles) {
    return guarded("nlhipBatchInfo", [&] {
        Batch* B = bat(b);
        if (nsys) *nsys = B->nsys;
        if (n) *n = B->n;
        if (nnz) *nnz = B->nnz;
        if (tiles) *tiles = B->ntiles;
    });
}

NLAPI int NLAPIENTRY nlhipBatchGetCRS(nlhip_batch b, uint32_t* row_off, uint32_t* rowptr,
                                      uint32_t* colind, double* val) {
    return guarded("nlhipBatchGetCRS", [&] {
        Batch* B = bat(b);
        std::lock_guard<std::mutex> lock(B->mu);
        hipStream_t s = B->stream.get();
        if (row_off)
            std::memcpy(row_off, B->h_row_off.data(), B->h_row_off.size() * 4);
        if (rowptr)
            NLHIP_CHECK(hipMemcpyAsync(rowptr, B->rowptr.get(), (size_t(B->n) + 1) * 4,
                                       hipMemcpyDefault, s));
        if (colind && B->nnz)
            NLHIP_CHECK(hipMemcpyAsync(colind, B->colind.get(), B->nnz * 4, hipMemcpyDefault, s));
        if (val && B->nnz)
            NLHIP_CHECK(hipMemcpyAsync(val, B->val.get(), B->nnz * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
    });
}

NLAPI int NLAPIENTRY nlhipBatchFallbacks(nlhip_batch b, uint32_t* n) {
    return guarded("nlhipBatchFallbacks", [&] {
        Batch* B = bat(b);
        std::lock_guard<std::mutex> lock(B->mu);
        if (n) *n = B->last_fallbacks;
    });
}

NLAPI int NLAPIENTRY nlhipBatchSetGraphs(nlhip_batch b, int on) {
    return guarded("nlhipBatchSetGraphs", [&] {
        Batch* B = bat(b);
        std::lock_guard<std::mutex> lock(B->mu);
        B->graphs = on != 0;
    });
}

NLAPI int NLAPIENTRY nlhipBatchSetLaunch(nlhip_batch b, uint32_t workgroups) {
    return guarded("nlhipBatchSetLaunch", [&] {
        Batch* B = bat(b);
        if (workgroups == 0 || workgroups > 65536)
            throw Error(NLHIP_ERR_INVALID, "workgroups must be in [1, 65536]");
        std::lock_guard<std::mutex> lock(B->mu);
        B->grid = workgroups;
        B->w64.graph.reset();
        B->w32.graph.reset();
    });
}

NLAPI int NLAPIENTRY nlhipBatchSetPersistent(nlhip_batch b, int on) {
    return guarded("nlhipBatchSetPersistent", [&] {
        Batch* B = bat(b);
        std::lock_guard<std::mutex> lock(B->mu);
        B->persistent = on != 0;
    });
}

NLAPI int NLAPIENTRY nlhipBatchSetSpmvPasses(nlhip_batch b, uint32_t passes) {
    return guarded("nlhipBatchSetSpmvPasses", [&] {
        Batch* B = bat(b);
        if (passes != 0 && passes != 1 && passes != 2 && passes != 4)
            throw Error(NLHIP_ERR_INVALID, "passes must be 0, 1, 2 or 4");
        std::lock_guard<std::mutex> lock(B->mu);
        B->spmv_passes = passes;
        set_spmv_runs(*B, B->w64.args);
        set_spmv_runs(*B, B->w32.args);
        B->w64.graph.reset();
        B->w32.graph.reset();
    });
}

// Autotuner: times `iters` FP64 iterations of the batched PCG from x = 0 on
// `rhs` for workgroup counts of 4, 16 and 64 per CU and for each SpMV run
// length (1, 2, 4 passes per wavefront, or chosen per launch by the active
// tiles), and keeps the fastest.  The first iterations have every system
// active, so a fixed run length must beat the adaptive one (and any shape
// the current one) by 2 % to be taken.  None of these changes a bit of the
// result.
NLAPI int NLAPIENTRY nlhipBatchAutotune(nlhip_batch b, const double* rhs, uint32_t iters,
                                        uint32_t* workgroups, uint32_t* passes,
                                        double* seconds) {
    return guarded("nlhipBatchAutotune", [&] {
        Batch* B = bat(b);
        if (!rhs) throw Error(NLHIP_ERR_INVALID, "null b");
        if (iters == 0) iters = 50;
        std::lock_guard<std::mutex> lock(B->mu);
        hipStream_t s = B->stream.get();
        copy_in(B->w64.b.get(), rhs, size_t(B->n) * 8, s);
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        p.eps = 0.0;  // run every iteration
        p.max_iter = iters;
        const uint32_t cu = uint32_t(std::max(1, device().cu_count));
        std::vector<uint32_t> grids;
        for (uint32_t m : {4u, 16u, 64u}) {
            const uint32_t g = uint32_t(std::max<uint64_t>(
                1, std::min<uint64_t>(uint64_t(cu) * m, B->ntiles)));
            if (std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
        }
        const uint32_t grid0 = B->grid, passes0 = B->spmv_passes;
        uint32_t best_g = grid0, best_p = passes0;
        double best = 1e300;
        std::vector<int32_t> its;
        std::vector<double> cgrel;
        std::vector<uint8_t> conv;
        std::vector<uint32_t> extra;
        // the current shape first, then the others
        std::vector<std::pair<uint32_t, uint32_t>> shapes{{grid0, passes0}};
        for (uint32_t g : grids)
            for (uint32_t ps : {0u, 1u, 2u, 4u})
                if (!(g == grid0 && ps == passes0)) shapes.emplace_back(g, ps);
        for (const auto& shape : shapes) {
            const uint32_t g = shape.first, ps = shape.second;
            B->grid = g;
            B->spmv_passes = ps;
            set_spmv_runs(*B, B->w64.args);
            set_spmv_runs(*B, B->w32.args);
            B->w64.graph.reset();
            B->w32.graph.reset();
            double t = 0.0;
            for (int rep = 0; rep < 2; ++rep) {  // the first captures the graphs
                NLHIP_CHECK(hipMemsetAsync(B->w64.x.get(), 0, size_t(B->n) * 8, s));
                Event e0, e1;
                e0.record(s);
                batch_solve_fp64(*B, p, its, cgrel, conv, extra);
                e1.record(s);
                e1.sync();
                t = 1e-3 * elapsed_ms(e0, e1);
            }
            if (best == 1e300 || t < 0.98 * best) {
                best = t;
                best_g = g;
                best_p = ps;
            }
        }
        B->grid = best_g;
        B->spmv_passes = best_p;
        set_spmv_runs(*B, B->w64.args);
        set_spmv_runs(*B, B->w32.args);
        B->w64.graph.reset();
        B->w32.graph.reset();
        if (workgroups) *workgroups = best_g;
        if (passes) *passes = best_p;
        if (seconds) *seconds = best / iters;
    });
}

// An atlas built on the device: the LSCM systems of many charts (one mesh
// whose vertices and triangles are grouped by chart) as one batch.
NLAPI int NLAPIENTRY nlhipBatchLSCMCreate(uint32_t nb_charts, const uint32_t* vertex_off,
                                          const double* points, uint32_t nb_triangles,
                                          const uint32_t* triangles, const uint32_t* pins,
                                          const double* pin_uv, uint32_t flags,
                                          nlhip_batch* out, double* rhs, uint32_t* n_out) {
    return guarded("nlhipBatchLSCMCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        if (nb_charts == 0 || !vertex_off || !points || !triangles || !pins || !pin_uv)
            throw Error(NLHIP_ERR_INVALID, "empty atlas");
        device();
        Stream stream;
        hipStream_t s = stream.get();
        // chart layout and pins (host copies; these arrays are small)
        std::vector<uint32_t> voff(size_t(nb_charts) + 1), pin(2 * size_t(nb_charts));
        std::vector<double> uv(4 * size_t(nb_charts));
        NLHIP_CHECK(hipMemcpyAsync(voff.data(), vertex_off, voff.size() * 4, hipMemcpyDefault, s));
        NLHIP_CHECK(hipMemcpyAsync(pin.data(), pins, pin.size() * 4, hipMemcpyDefault, s));
        NLHIP_CHECK(hipMemcpyAsync(uv.data(), pin_uv, uv.size() * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        if (voff[0] != 0) throw Error(NLHIP_ERR_INVALID, "vertex_off[0] != 0");
        const uint32_t nv = voff[nb_charts];
        if (uint64_t(nv) * 2 > 0xffffffffull) throw Error(NLHIP_ERR_UNSUPPORTED, "too many vertices");
        std::vector<uint32_t> row_off(size_t(nb_charts) + 1, 0);
        std::vector<uint32_t> free_index(2 * size_t(nv));
        std::vector<double> values(2 * size_t(nv), 0.0);
        std::vector<uint8_t> locked(nv, 0);
        for (uint32_t c = 0; c < nb_charts; ++c) {
            const uint32_t lo = voff[c], hi = voff[c + 1];
            if (hi < lo + 3) throw Error(NLHIP_ERR_INVALID, "a chart needs at least 3 vertices");
            const uint32_t p0 = pin[2 * c], p1 = pin[2 * c + 1];
            if (p0 == p1 || p0 < lo || p0 >= hi || p1 < lo || p1 >= hi)
                throw Error(NLHIP_ERR_INVALID, "pins must be two distinct vertices of their chart");
            for (int k = 0; k < 4; ++k)
                if (!std::isfinite(uv[4 * c + k]))
                    throw Error(NLHIP_ERR_NONFINITE, "pin position is NaN or Inf");
            locked[p0] = locked[p1] = 1;
            values[2 * size_t(p0)] = uv[4 * c];
            values[2 * size_t(p0) + 1] = uv[4 * c + 1];
            values[2 * size_t(p1)] = uv[4 * c + 2];
            values[2 * size_t(p1) + 1] = uv[4 * c + 3];
            row_off[c + 1] = row_off[c] + 2 * (hi - lo - 2);
        }
        uint32_t f = 0;
        for (uint32_t v = 0; v < nv; ++v)
            for (int k = 0; k < 2; ++k) free_index[2 * size_t(v) + k] = locked[v] ? 0xffffffffu : f++;
        // rows and normal equations on the device
        DevBuf<double> d_p;
        DevBuf<uint32_t> d_t, d_voff(voff.size()), d_free(free_index.size());
        DevBuf<double> d_values(values.size());
        const double* P = points;
        const uint32_t* T = triangles;
        if (!is_device_pointer(points)) {
            d_p.alloc(3 * size_t(nv));
            copy_in(d_p.get(), points, 3 * size_t(nv) * 8, s);
            P = d_p.get();
        }
        if (!is_device_pointer(triangles)) {
            d_t.alloc(3 * size_t(nb_triangles ? nb_triangles : 1));
            copy_in(d_t.get(), triangles, 3 * size_t(nb_triangles) * 4, s);
            T = d_t.get();
        }
        copy_in(d_voff.get(), voff.data(), voff.size() * 4, s);
        copy_in(d_free.get(), free_index.data(), free_index.size() * 4, s);
        copy_in(d_values.get(), values.data(), values.size() * 8, s);
        LSRows rows;
        lscm_rows_device(nv, P, nb_triangles, T, rows, s);   // checks T < nv, finite P
        check_triangle_charts(nb_triangles, T, d_voff.get(), nb_charts, s);
        LSInput in;
        in.nrows = rows.nrows;
        in.nvars = 2 * nv;
        in.nsys = 1;
        in.nfree = f;
        in.nent = 6ull * rows.nrows;
        in.rp = rows.rp.get();
        in.var = rows.var.get();
        in.coef = rows.coef.get();
        in.free_index = d_free.get();
        in.values = d_values.get();
        LSOutput o;
        ls_assemble_device(in, o, s);
        if (rhs) {
            NLHIP_CHECK(hipMemcpyAsync(rhs, o.rhs.get(), size_t(f) * 8, hipMemcpyDefault, s));
            NLHIP_CHECK(hipStreamSynchronize(s));
        }
        if (n_out) *n_out = f;
        Batch* B = batch_adopt(nb_charts, row_off.data(), o.nnz, std::move(o.rowptr),
                               std::move(o.colind), std::move(o.val), flags);
        *out = reinterpret_cast<nlhip_batch>(B);
    });
}

NLAPI int NLAPIENTRY nlhipBatchSolve(nlhip_batch b, const double* rhs, double* x,
                                     const nlhip_solve_params* params,
                                     nlhip_solve_stats* per_system, nlhip_solve_stats* total) {
    return guarded("nlhipBatchSolve", [&] {
        Batch* B = bat(b);
        if (!rhs || !x) throw Error(NLHIP_ERR_INVALID, "null b or x");
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        if (params) p = *params;
        if (p.precond != NL_PRECOND_JACOBI && p.precond != NL_PRECOND_NONE)
            throw Error(NLHIP_ERR_UNSUPPORTED, "only Jacobi or no preconditioner on the GPU");
        std::lock_guard<std::mutex> lock(B->mu);
        if (take_injected_failure()) throw Error(NLHIP_ERR_INJECTED, "injected failure");
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = B->stream.get();
        copy_in(B->w64.b.get(), rhs, size_t(B->n) * 8, s);
        copy_in(B->w64.x.get(), x, size_t(B->n) * 8, s);
        Event e0, e1;
        e0.record(s);
        std::vector<int32_t> its;
        std::vector<double> cgrel, rel;
        std::vector<uint8_t> conv;
        std::vector<uint32_t> outer, extra;
        if (p.precision == NLHIP_PREC_FP64) {
            batch_solve_fp64(*B, p, its, cgrel, conv, extra);
        } else if (p.precision == NLHIP_PREC_MIXED) {
            batch_solve_mixed(*B, p, its, cgrel, conv, outer);
        } else if (p.precision == NLHIP_PREC_MIXED_FALLBACK) {
            uint32_t nfb = 0;
            batch_solve_mixed_fallback(*B, p, its, cgrel, conv, outer, extra, nfb);
            B->last_fallbacks = nfb;
        } else {
            throw Error(NLHIP_ERR_INVALID, "unknown precision");
        }
        e1.record(s);
        e1.sync();
        const double dev_s = 1e-3 * elapsed_ms(e0, e1);
        true_residuals(*B, rel);
        copy_out(x, B->w64.x.get(), size_t(B->n) * 8, s);
        B->stream.sync();
        const double wall_s =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fill_stats(*B, its, rel, cgrel, conv, outer.empty() ? nullptr : &outer, extra, per_system,
                   total, dev_s, wall_s);
    });
}

}  // extern "C"
