// Solve server: admits requests (one system, one or more right-hand sides)
// under an HBM byte budget and a slot limit, runs the admitted ones
// concurrently -- each on its own resident system and HIP stream, so their
// kernels overlap on the device -- and hands back solutions with per-request
// timestamps for latency accounting.
//
// Admission is first-come first-served: the oldest waiting request is
// admitted as soon as a slot is free and its device footprint fits in what
// the budget has left; a request larger than the whole budget is rejected
// with NLHIP_ERR_BUDGET instead of queueing forever.  Every request's kernels
// depend only on that request (its own matrix, vectors, launch shape), so a
// result is bitwise the same whether it ran alone or next to others.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>

#include "nlhip_internal.h"

namespace nlhip {

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    uint32_t n = 0, k = 1;
    std::vector<uint32_t> rowptr, colind;
    std::vector<double> val, B, X;
    nlhip_solve_params params{};
    uint64_t bytes = 0;
    int status = NLHIP_OK;
    bool finished = false;
    nlhip_solve_stats stats{};
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// Device bytes a request holds while it runs: matrix, Jacobi diagonal, the
// FP64 CG workspace (or the n x k multi right-hand-side workspace), FP32
// copies for mixed precision, and partial-sum buffers.
uint64_t footprint(uint32_t n, uint64_t nnz, uint32_t k, bool mixed) {
    const uint64_t parts = 4ull * 8 * 1024;                       // per-workgroup partials
    uint64_t b = 4ull * (n + 1) + 12ull * nnz + 8ull * n;        // rowptr, colind, val, dinv
    b += 2ull * nnz + 4ull * (nnz / 1024 + 2);                    // 16-bit column offsets
    b += 8ull * nnz + nnz / 2 + 2ull * n + 4ull * (nnz / 4096 + 2); // 2x2 block form
    if (k <= 1) b += 8ull * 8 * n + parts;                        // x b r h d q + scratch
    else b += 7ull * 8 * n * k + parts * k;                       // interleaved vectors
    if (mixed) b += 8ull * nnz + 4ull * n + (k <= 1 ? 7ull * 4 * n : 6ull * 4 * n * k);
    return b + (256ull << 10);                                     // allocation slack
}

}  // namespace

struct Server {
    uint64_t budget = 0;
    uint32_t slots = 1;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    uint64_t in_use = 0, peak_in_use = 0;
    uint32_t running = 0, peak_running = 0;
    uint64_t completed = 0, failed = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void worker() {
        for (;;) {
            std::shared_ptr<Request> R;
            {
                std::unique_lock<std::mutex> lock(mu);
                auto admissible = [&] {
                    return !pending.empty() && running < slots &&
                           in_use + pending.front()->bytes <= budget;
                };
                cv_work.wait(lock, [&] { return (stop && pending.empty()) || admissible(); });
                if (!admissible()) return;  // stopping and nothing left to run
                R = pending.front();
                pending.pop_front();
                in_use += R->bytes;
                peak_in_use = std::max(peak_in_use, in_use);
                ++running;
                peak_running = std::max(peak_running, running);
                R->t_admit = now();
            }
            run(*R);
            {
                std::lock_guard<std::mutex> lock(mu);
                in_use -= R->bytes;
                --running;
                R->t_done = now();
                R->finished = true;
                (R->status == NLHIP_OK ? completed : failed) += 1;
                // inputs are no longer needed; keep only the solution
                R->rowptr = {};
                R->colind = {};
                R->val = {};
                R->B = {};
            }
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }

    void run(Request& R) {
        R.status = guarded("nlhip server request", [&] {
            std::unique_ptr<System> S(system_create(R.n, R.rowptr.data(), R.colind.data(),
                                                    R.val.data(), 0));
            if (R.k <= 1) {
                solve(*S, R.B.data(), R.X.data(), R.params, R.stats);
            } else {
                solve_multi(*S, R.k, R.B.data(), R.X.data(), R.params, nullptr, &R.stats);
            }
        });
        R.stats.status = R.status;
    }
};

}  // namespace nlhip

using namespace nlhip;

static Server* srv(nlhip_server h) {
    if (!h) throw Error(NLHIP_ERR_INVALID, "null server handle");
    return reinterpret_cast<Server*>(h);
}

extern "C" {

NLAPI void NLAPIENTRY nlhipServerConfigDefault(nlhip_server_config* c) {
    if (!c) return;
    c->hbm_budget_bytes = 200ull << 30;
    c->max_active = 8;
}

NLAPI int NLAPIENTRY nlhipServerCreate(const nlhip_server_config* config, nlhip_server* out) {
    return guarded("nlhipServerCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        nlhip_server_config c;
        nlhipServerConfigDefault(&c);
        if (config) c = *config;
        if (c.max_active == 0 || c.max_active > 64)
            throw Error(NLHIP_ERR_INVALID, "max_active must be in [1, 64]");
        const Device& d = device();
        auto S = std::make_unique<Server>();
        S->budget = std::min<uint64_t>(c.hbm_budget_bytes, d.hbm_bytes);
        S->slots = c.max_active;
        for (uint32_t i = 0; i < S->slots; ++i) S->workers.emplace_back([p = S.get()] { p->worker(); });
        *out = reinterpret_cast<nlhip_server>(S.release());
    });
}

NLAPI int NLAPIENTRY nlhipServerSubmit(nlhip_server h, uint32_t n, const uint32_t* rowptr,
                                       const uint32_t* colind, const double* val, uint32_t k,
                                       const double* B, const double* X0,
                                       const nlhip_solve_params* params, uint64_t* id) {
    return guarded("nlhipServerSubmit", [&] {
        Server* S = srv(h);
        if (!id) throw Error(NLHIP_ERR_INVALID, "null id");
        *id = 0;
        if (n == 0 || !rowptr || !B) throw Error(NLHIP_ERR_INVALID, "empty request");
        if (k == 0 || k > 32) throw Error(NLHIP_ERR_UNSUPPORTED, "1 to 32 right-hand sides");
        auto R = std::make_shared<Request>();
        R->n = n;
        R->k = k;
        const uint64_t nnz = rowptr[n];
        R->rowptr.assign(rowptr, rowptr + n + 1);
        R->colind.assign(colind, colind + nnz);
        R->val.assign(val, val + nnz);
        R->B.assign(B, B + size_t(n) * k);
        if (X0) R->X.assign(X0, X0 + size_t(n) * k);
        else R->X.assign(size_t(n) * k, 0.0);
        nlhipSolveParamsDefault(&R->params);
        if (params) R->params = *params;
        R->bytes = footprint(n, nnz, k, R->params.precision == NLHIP_PREC_MIXED);
        std::lock_guard<std::mutex> lock(S->mu);
        if (S->stop) throw Error(NLHIP_ERR_INVALID, "server is shutting down");
        R->id = S->next_id++;
        R->t_submit = S->now();
        S->all[R->id] = R;
        *id = R->id;
        if (R->bytes > S->budget) {
            R->status = NLHIP_ERR_BUDGET;
            R->stats.status = R->status;
            R->finished = true;
            R->t_admit = R->t_done = R->t_submit;
            S->failed += 1;
            set_last_error("request needs " + std::to_string(R->bytes) +
                           " device bytes, more than the whole budget");
        } else {
            S->pending.push_back(R);
        }
        S->cv_work.notify_all();
        S->cv_done.notify_all();
    });
}

NLAPI int NLAPIENTRY nlhipServerWait(nlhip_server h, uint64_t id, double* X,
                                     nlhip_request_result* result, double timeout_seconds) {
    return guarded("nlhipServerWait", [&] {
        Server* S = srv(h);
        std::unique_lock<std::mutex> lock(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) throw Error(NLHIP_ERR_INVALID, "unknown request id");
        std::shared_ptr<Request> R = it->second;
        auto pred = [&] { return R->finished; };
        if (timeout_seconds > 0) {
            if (!S->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_seconds), pred))
                throw Error(NLHIP_ERR_INVALID, "request not finished within the timeout");
        } else {
            S->cv_done.wait(lock, pred);
        }
        if (X && R->status == NLHIP_OK) std::memcpy(X, R->X.data(), R->X.size() * sizeof(double));
        if (result) {
            result->id = R->id;
            result->status = R->status;
            result->k = R->k;
            result->n = R->n;
            result->device_bytes = R->bytes;
            result->stats = R->stats;
            result->t_submit = R->t_submit;
            result->t_admit = R->t_admit;
            result->t_done = R->t_done;
        }
        S->all.erase(it);
        if (R->status != NLHIP_OK) throw Error(R->status, "request failed: status " +
                                                              std::to_string(R->status));
    });
}

NLAPI int NLAPIENTRY nlhipServerInfo(nlhip_server h, nlhip_server_info* info) {
    return guarded("nlhipServerInfo", [&] {
        Server* S = srv(h);
        if (!info) throw Error(NLHIP_ERR_INVALID, "null info");
        std::lock_guard<std::mutex> lock(S->mu);
        info->budget_bytes = S->budget;
        info->in_use_bytes = S->in_use;
        info->peak_in_use_bytes = S->peak_in_use;
        info->pending = uint32_t(S->pending.size());
        info->running = S->running;
        info->peak_running = S->peak_running;
        info->completed = S->completed;
        info->failed = S->failed;
    });
}

NLAPI int NLAPIENTRY nlhipServerDestroy(nlhip_server h) {
    return guarded("nlhipServerDestroy", [&] {
        if (!h) return;
        Server* S = srv(h);
        {
            std::lock_guard<std::mutex> lock(S->mu);
            S->stop = true;
        }
        S->cv_work.notify_all();
        for (auto& t : S->workers) t.join();
        delete S;
    });
}

}  // extern "C"
