// Solve server: admits requests (one system, one or more right-hand sides)
// under an HBM byte budget and a slot limit, runs the admitted ones
// concurrently -- each solve on its own HIP stream, so their kernels overlap
// on the device -- and hands back solutions with per-request timestamps for
// latency accounting.
//
// Admission is first-come first-served: the oldest waiting request is
// admitted as soon as a slot is free and its device footprint fits in what
// the budget has left; a request larger than the whole budget is rejected
// with NLHIP_ERR_BUDGET instead of queueing forever.
//
// Batching: a small single-right-hand-side request (n <= kBatchRows) is
// solved by the batched solver, together with the other waiting small
// requests that have the same solve parameters (taken in arrival order,
// within the budget and kBatchSystems / kBatchTotalRows); a batch takes one
// slot.  Such a request always goes this way, alone as a batch of one, and
// a batched system's result does not depend on the rest of its batch, so
// every result is bitwise the same whether it ran alone or next to others.
// Larger and multi-right-hand-side requests run on their own resident
// system.  Validation and the fault-injection hook act on each request
// before batching, so a failing request fails alone.
//
// Resident systems: a request can name a system the caller keeps resident
// (nlhipServerSubmitSystem): only its right-hand sides travel, nothing of the
// matrix is uploaded, and it is charged only its solve workspace.  Requests
// on one system run one at a time in arrival order (the system has one
// workspace and one stream); while its system is busy, later requests may
// be admitted before it.  A request on a resident system is never batched,
// so its result is bitwise that of nlhipSolve / nlhipSolveMulti on the
// system, alone or in any mix.
#include <algorithm>
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

constexpr uint32_t kBatchRows = 1u << 18;        // largest request that is batched
constexpr uint32_t kBatchSystems = 256;          // requests per batch
constexpr uint64_t kBatchTotalRows = 1ull << 23; // rows per batch

struct Request {
    uint64_t id = 0;
    uint32_t n = 0, k = 1;
    bool batchable = false;
    System* resident = nullptr;   // a caller's resident system (no upload)
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
// Device bytes of the solve workspace alone (a request on a resident system).
uint64_t workspace_footprint(uint32_t n, uint32_t k, bool mixed) {
    const uint64_t parts = 4ull * 8 * 1024;
    uint64_t b = k <= 1 ? 8ull * 8 * n + parts : 7ull * 8 * n * k + parts * k;
    if (mixed) b += k <= 1 ? 7ull * 4 * n : 6ull * 4 * n * k;
    return b + (64ull << 10);
}

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
    std::map<System*, uint32_t> busy;  // resident systems with a request running
    uint64_t next_id = 1;
    uint64_t in_use = 0, peak_in_use = 0;
    uint32_t running = 0, peak_running = 0;  // requests on the device
    uint32_t solves = 0;                     // solves (single or batched) in flight
    uint64_t completed = 0, failed = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void worker() {
        for (;;) {
            std::vector<std::shared_ptr<Request>> group;
            uint64_t bytes = 0;
            {
                std::unique_lock<std::mutex> lock(mu);
                // the oldest waiting request whose bytes fit; a request on a
                // busy resident system is passed over until that system is free
                auto pick = [&]() -> std::deque<std::shared_ptr<Request>>::iterator {
                    if (solves >= slots) return pending.end();
                    for (auto it = pending.begin(); it != pending.end(); ++it) {
                        if ((*it)->resident && busy.count((*it)->resident)) continue;
                        return in_use + (*it)->bytes <= budget ? it : pending.end();
                    }
                    return pending.end();
                };
                auto chosen = pending.end();
                cv_work.wait(lock, [&] {
                    chosen = pick();
                    return (stop && pending.empty()) || chosen != pending.end();
                });
                if (chosen == pending.end()) return;  // stopping and nothing left to run
                group.push_back(*chosen);
                pending.erase(chosen);
                bytes = group[0]->bytes;
                if (group[0]->resident) busy[group[0]->resident] += 1;
                if (group[0]->batchable) {
                    // compatible waiting requests, in arrival order
                    uint64_t rows = group[0]->n;
                    for (auto it = pending.begin();
                         it != pending.end() && group.size() < kBatchSystems;) {
                        const Request& C = **it;
                        if (C.batchable && rows + C.n <= kBatchTotalRows &&
                            in_use + bytes + C.bytes <= budget &&
                            std::memcmp(&C.params, &group[0]->params, sizeof(C.params)) == 0) {
                            rows += C.n;
                            bytes += C.bytes;
                            group.push_back(*it);
                            it = pending.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
                in_use += bytes;
                peak_in_use = std::max(peak_in_use, in_use);
                ++solves;
                running += uint32_t(group.size());
                peak_running = std::max(peak_running, running);
                const double t = now();
                for (auto& R : group) R->t_admit = t;
            }
            if (group[0]->batchable) run_batch(group);
            else run(*group[0]);
            {
                std::lock_guard<std::mutex> lock(mu);
                in_use -= bytes;
                --solves;
                running -= uint32_t(group.size());
                if (System* Rs = group[0]->resident) {
                    if (--busy[Rs] == 0) busy.erase(Rs);
                }
                const double t = now();
                for (auto& R : group) {
                    R->t_done = t;
                    R->finished = true;
                    (R->status == NLHIP_OK ? completed : failed) += 1;
                    // inputs are no longer needed; keep only the solution
                    R->rowptr = {};
                    R->colind = {};
                    R->val = {};
                    R->B = {};
                }
            }
            cv_work.notify_all();
            cv_done.notify_all();
        }
    }

    void run(Request& R) {
        R.status = guarded("nlhip server request", [&] {
            std::unique_ptr<System> own;
            System* S = R.resident;
            if (!S) {
                own.reset(system_create(R.n, R.rowptr.data(), R.colind.data(), R.val.data(), 0));
                S = own.get();
            }
            if (R.k <= 1) {
                solve(*S, R.B.data(), R.X.data(), R.params, R.stats);
            } else {
                solve_multi(*S, R.k, R.B.data(), R.X.data(), R.params, nullptr, &R.stats);
            }
        });
        R.stats.status = R.status;
    }

    // Validates every request (a failing one fails alone), then solves the
    // others as one batch and hands each its own solution and statistics.
    void run_batch(std::vector<std::shared_ptr<Request>>& group) {
        std::vector<Request*> ok;
        for (auto& R : group) {
            R->status = guarded("nlhip server request", [&] {
                if (take_injected_failure())
                    throw Error(NLHIP_ERR_INJECTED, "injected failure");
                validate(*R);
            });
            R->stats.status = R->status;
            if (R->status == NLHIP_OK) ok.push_back(R.get());
        }
        if (ok.empty()) return;
        const uint32_t ns = uint32_t(ok.size());
        std::vector<uint32_t> row_off(size_t(ns) + 1, 0), rowptr, colind;
        std::vector<double> val, b, x;
        for (uint32_t i = 0; i < ns; ++i) row_off[i + 1] = row_off[i] + ok[i]->n;
        rowptr.reserve(size_t(row_off[ns]) + 1);
        uint64_t nnz = 0;
        for (Request* R : ok) {
            for (uint32_t r = 0; r < R->n; ++r) rowptr.push_back(uint32_t(nnz + R->rowptr[r]));
            nnz += R->rowptr[R->n];
            if (nnz > 0xffffffffull) break;
        }
        rowptr.push_back(uint32_t(nnz));
        colind.reserve(nnz);
        val.reserve(nnz);
        for (uint32_t i = 0; i < ns; ++i) {
            for (uint32_t c : ok[i]->colind) colind.push_back(c + row_off[i]);
            val.insert(val.end(), ok[i]->val.begin(), ok[i]->val.end());
            b.insert(b.end(), ok[i]->B.begin(), ok[i]->B.end());
            x.insert(x.end(), ok[i]->X.begin(), ok[i]->X.end());
        }
        std::vector<nlhip_solve_stats> per(ns);
        nlhip_batch Bh = nullptr;
        const nlhip_solve_params& p = ok[0]->params;
        int rc = nnz > 0xffffffffull ? NLHIP_ERR_UNSUPPORTED : NLHIP_OK;
        if (rc == NLHIP_OK)
            rc = nlhipBatchCreate(ns, row_off.data(), rowptr.data(), colind.data(), val.data(),
                                  p.precision != NLHIP_PREC_FP64 ? NLHIP_SYSTEM_FP32 : 0u, &Bh);
        if (rc == NLHIP_OK) rc = nlhipBatchSolve(Bh, b.data(), x.data(), &p, per.data(), nullptr);
        if (Bh) nlhipBatchDestroy(Bh);
        for (uint32_t i = 0; i < ns; ++i) {
            Request& R = *ok[i];
            R.status = rc;
            if (rc == NLHIP_OK) {
                R.stats = per[i];
                std::copy(x.begin() + row_off[i], x.begin() + row_off[i + 1], R.X.begin());
            }
            R.stats.status = rc;
        }
    }

    // Registers a new request: queued, or failed at once when it can never
    // fit the budget.
    void enqueue(const std::shared_ptr<Request>& R, uint64_t* id) {
        std::lock_guard<std::mutex> lock(mu);
        if (stop) throw Error(NLHIP_ERR_INVALID, "server is shutting down");
        R->id = next_id++;
        R->t_submit = now();
        all[R->id] = R;
        *id = R->id;
        if (R->bytes > budget) {
            R->status = NLHIP_ERR_BUDGET;
            R->stats.status = R->status;
            R->finished = true;
            R->t_admit = R->t_done = R->t_submit;
            failed += 1;
            set_last_error("request needs " + std::to_string(R->bytes) +
                           " device bytes, more than the whole budget");
        } else {
            pending.push_back(R);
        }
        cv_work.notify_all();
        cv_done.notify_all();
    }

    // The checks system creation makes, per request: CRS structure, finite
    // values and right-hand side.
    static void validate(const Request& R) {
        if (R.rowptr.empty() || R.rowptr[0] != 0) throw Error(NLHIP_ERR_INVALID, "rowptr[0] != 0");
        for (uint32_t i = 0; i < R.n; ++i)
            if (R.rowptr[i + 1] < R.rowptr[i])
                throw Error(NLHIP_ERR_INVALID, "rowptr is not monotone");
        for (uint32_t c : R.colind)
            if (c >= R.n) throw Error(NLHIP_ERR_INVALID, "colind out of range");
        for (double v : R.val)
            if (!std::isfinite(v)) throw Error(NLHIP_ERR_NONFINITE, "matrix value is NaN or Inf");
        for (double v : R.B)
            if (!std::isfinite(v))
                throw Error(NLHIP_ERR_NONFINITE, "right-hand side is NaN or Inf");
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
        R->bytes = footprint(n, nnz, k, R->params.precision != NLHIP_PREC_FP64);
        R->batchable = (k == 1 && n <= kBatchRows);
        S->enqueue(R, id);
    });
}

NLAPI int NLAPIENTRY nlhipServerSubmitSystem(nlhip_server h, nlhip_system sys, uint32_t k,
                                             const double* B, const double* X0,
                                             const nlhip_solve_params* params, uint64_t* id) {
    return guarded("nlhipServerSubmitSystem", [&] {
        Server* S = srv(h);
        if (!id) throw Error(NLHIP_ERR_INVALID, "null id");
        *id = 0;
        if (!sys) throw Error(NLHIP_ERR_INVALID, "null system handle");
        if (!B) throw Error(NLHIP_ERR_INVALID, "null right-hand sides");
        if (k == 0 || k > 32) throw Error(NLHIP_ERR_UNSUPPORTED, "1 to 32 right-hand sides");
        System* Sys = reinterpret_cast<System*>(sys);
        const uint32_t n = Sys->n;
        auto R = std::make_shared<Request>();
        R->n = n;
        R->k = k;
        R->resident = Sys;
        R->B.assign(B, B + size_t(n) * k);
        for (double v : R->B)
            if (!std::isfinite(v))
                throw Error(NLHIP_ERR_NONFINITE, "right-hand side is NaN or Inf");
        if (X0) R->X.assign(X0, X0 + size_t(n) * k);
        else R->X.assign(size_t(n) * k, 0.0);
        nlhipSolveParamsDefault(&R->params);
        if (params) R->params = *params;
        R->bytes = workspace_footprint(n, k, R->params.precision != NLHIP_PREC_FP64);
        S->enqueue(R, id);
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
