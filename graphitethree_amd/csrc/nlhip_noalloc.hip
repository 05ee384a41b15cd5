// nlhipSolveNoAlloc: the FP64 PCG of nlhipSolve on a resident system for
// callers that cannot allocate (tight serving loops, code that must not call
// hipMalloc).  Every work buffer of the solve -- the staged b and x, the
// recurrence's vectors r, h, d, q, the per-workgroup partials, the residual
// partials and the status block -- lives in `scratch`, one fixed device
// array of this file (104 MiB, held by each GPU the library runs on until
// the process ends).  A call makes no hipMalloc, hipFree, hipHostMalloc,
// event creation or graph instantiation: it launches the same kernels as
// nlhipSolve directly (the graph path replays the same launches), polls the
// status into host stack memory, and drives them with the same stopping and
// confirmation logic (pcg_drive), so its result is bitwise nlhipSolve's.
//
// Every call on one GPU shares that GPU's array: two calls that overlap on
// one GPU (from two threads, on any systems) race and corrupt each other.
// Nothing in the array has to start at zero: the start launches write the
// residual, the partials and the whole status block before anything reads
// them.
#include <chrono>
#include <cstring>
#include <mutex>

#include <hip/hip_runtime.h>

#include "nlhip_internal.h"
#include "nlhip_kernels.h"

namespace nlhip {
namespace {

constexpr uint32_t kNoAllocMaxRows = NLHIP_NOALLOC_MAX_ROWS;  // 2,097,152 rows
constexpr uint32_t kNoAllocMaxGrid = 65536;      // nlhipSystemSetLaunch's limit
constexpr size_t kVec = kNoAllocMaxRows;         // doubles per vector (16 MiB)
constexpr size_t kPart = size_t(kSpmvGridMul + 3) * kNoAllocMaxGrid;
constexpr size_t kRes = 2 * size_t(kNoAllocMaxGrid) + 8;  // residual partials + 4 norms
constexpr size_t kStatus = 32;                   // doubles reserved for CGStatus<double>
constexpr size_t kScratchDoubles = 6 * kVec + kPart + kRes + kStatus;
static_assert(sizeof(CGStatus<double>) <= kStatus * sizeof(double), "status block");
static_assert(alignof(CGStatus<double>) <= 8, "status block alignment");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

}  // namespace

// The fixed work array.  In host code `scratch` names a host-side stand-in,
// not device memory: only the address returned by hipGetSymbolAddress (one
// per device, scratch_for_current_device) is passed to kernels and copies.
static __device__ __align__(256) double scratch[kScratchDoubles];

// The current device's copy of `scratch`, looked up once per device (the
// device that device() made current).
static double* scratch_for_current_device() {
    int dev = 0;
    NLHIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(NLHIP_ERR_INVALID, "device id out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            throw Error(NLHIP_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                           hipGetErrorString(e));
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

namespace {

// The device loop run with direct launches; status polls into stack memory.
struct DirectPcgOps final : PcgOps {
    System& S;
    const CGArgs<double>& a;
    int code;
    bool precond;
    double* res;  // residual partials (2 x grid) and the 4 norms after them
    DirectPcgOps(System& s, const CGArgs<double>& args, int c, bool pc, double* r)
        : S(s), a(args), code(c), precond(pc), res(r) {}

    CGStatus<double> status() {
        CGStatus<double> st;
        NLHIP_CHECK(hipMemcpyAsync(&st, a.st, sizeof(st), hipMemcpyDeviceToHost, S.stream.get()));
        S.stream.sync();
        return st;
    }
    CGStatus<double> start(double eps) override {
        NLHIP_CHECK(cg_start<double>(a, code, precond, eps, 0, S.stream.get()));
        return status();
    }
    void reopen(const CGStatus<double>& cur, double thresh, int cap) override {
        CGStatus<double> st = cur;
        st.stop = 0;
        st.done = 0;
        st.alpha = 0.0;  // the stopping direction launch applied x -= alpha d already
        st.thresh = thresh;
        st.max_iter = cap;
        NLHIP_CHECK(hipMemcpyAsync(a.st, &st, sizeof(st), hipMemcpyHostToDevice, S.stream.get()));
        S.stream.sync();
    }
    CGStatus<double> run(uint32_t K, uint64_t budget) override {
        hipStream_t s = S.stream.get();
        if (K == 0) K = 1;
        // the device stops by itself at its cap; past that the status must say so
        const uint64_t limit = budget + 2ull * K + 8;
        uint64_t covered = 0;
        for (;;) {
            for (uint32_t i = 0; i < K; ++i) NLHIP_CHECK(cg_step<double>(a, code, precond, s));
            covered += K;
            const CGStatus<double> st = status();
            if (st.done) return st;
            if (covered > limit) throw Error(NLHIP_ERR_HIP, "CG did not reach its stopping test");
        }
    }
    void norms(double& rr, double& bb) override {
        hipStream_t s = S.stream.get();
        const uint32_t P = S.grid;
        const SpmvOps<double> o = S.ops<double>();
        double* out = res + 2 * size_t(P);
        NLHIP_CHECK(residual64(S.n, o.rp, o.cols, o.val, a.x, a.b, nullptr, res, res + P, o.code,
                               P, s));
        NLHIP_CHECK(reduce_final(res, P, out, s));
        NLHIP_CHECK(reduce_final(res + P, P, out + 2, s));
        double h[4];
        NLHIP_CHECK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
        S.stream.sync();
        rr = h[0];
        bb = h[2];
    }
};

void solve_noalloc(System& S, const double* b, double* x, const nlhip_solve_params& p,
                   nlhip_solve_stats& st) {
    std::lock_guard<std::mutex> lock(S.mu);
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(&st, 0, sizeof(st));
    if (take_injected_failure()) throw Error(NLHIP_ERR_INJECTED, "injected failure");
    if (p.precision != NLHIP_PREC_FP64)
        throw Error(NLHIP_ERR_UNSUPPORTED, "nlhipSolveNoAlloc runs FP64 CG only");
    if (p.precond != NL_PRECOND_JACOBI && p.precond != NL_PRECOND_NONE)
        throw Error(NLHIP_ERR_UNSUPPORTED, "only Jacobi or no preconditioner on the GPU");
    if (S.n > kNoAllocMaxRows)
        throw Error(NLHIP_ERR_INVALID,
                    "nlhipSolveNoAlloc: " + std::to_string(S.n) + " rows exceed its fixed work "
                    "array of " + std::to_string(kNoAllocMaxRows) + " rows; use nlhipSolve");
    if (S.grid == 0 || S.grid > kNoAllocMaxGrid)
        throw Error(NLHIP_ERR_INVALID, "nlhipSolveNoAlloc: launch grid out of range");
    double* base = scratch_for_current_device();
    CGBuffers<double> w;
    w.x = base;
    w.b = base + kVec;
    w.r = base + 2 * kVec;
    w.h = base + 3 * kVec;
    w.d = base + 4 * kVec;
    w.q = base + 5 * kVec;
    w.part = base + 6 * kVec;
    double* res = w.part + kPart;
    w.st = reinterpret_cast<CGStatus<double>*>(res + kRes);
    const bool precond = p.precond == NL_PRECOND_JACOBI;
    CGArgs<double> a{};
    int code = 0;
    fill_cg_args<double>(S, w, precond ? S.dinv.get() : nullptr, a, code);
    hipStream_t s = S.stream.get();
    copy_in(w.b, b, size_t(S.n) * 8, s);
    copy_in(w.x, x, size_t(S.n) * 8, s);
    DirectPcgOps ops(S, a, code, precond, res);
    PcgState P;
    pcg_open(P, ops, p, precond, p.check_every ? p.check_every : auto_check_every(S));
    pcg_drive(P, ops, p.max_iter);
    pcg_stats(P, st);
    if (P.rr >= 0.0 && P.bb > 0.0 && P.p.confirm) {
        st.rel_residual = std::sqrt(P.rr / P.bb);
    } else {
        double rr = 0, bb = 0;
        ops.norms(rr, bb);
        st.rel_residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    }
    copy_out(x, w.x, size_t(S.n) * 8, s);
    S.stream.sync();
    st.device_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st.wall_seconds = st.device_seconds;
    st.status = NLHIP_OK;
}

}  // namespace
}  // namespace nlhip

using namespace nlhip;

static System* sys(nlhip_system h) {
    if (!h) throw Error(NLHIP_ERR_INVALID, "null system handle");
    return reinterpret_cast<System*>(h);
}

NLAPI int NLAPIENTRY nlhipSolveNoAlloc(nlhip_system s, const double* b, double* x,
                                       const nlhip_solve_params* params,
                                       nlhip_solve_stats* stats) {
    nlhip_solve_stats local;
    nlhip_solve_stats& st = stats ? *stats : local;
    const int rc = guarded("nlhipSolveNoAlloc", [&] {
        System* S = sys(s);
        if (!b || !x) throw Error(NLHIP_ERR_INVALID, "null b or x");
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        if (params) p = *params;
        solve_noalloc(*S, b, x, p, st);
    });
    if (rc != NLHIP_OK) st.status = rc;
    return rc;
}
