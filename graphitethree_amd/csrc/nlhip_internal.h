// Host-side runtime of libnlhip.so: device handle, buffers, resident
// systems and the solve drivers.  C++17, compiled with g++ against the HIP
// runtime API; kernels are reached through nlhip_kernels.h.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <shared_mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "nl.h"
#include "nl_hip.h"
#include "nlhip_kernels.h"

namespace nlhip {

// ------------------------------------------------------------------ errors

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* expr, const char* file, int line);

#define NLHIP_CHECK(expr)                                                   \
    do {                                                                    \
        hipError_t nlhip_rc_ = (expr);                                      \
        if (nlhip_rc_ != hipSuccess) ::nlhip::throw_hip(nlhip_rc_, #expr, __FILE__, __LINE__); \
    } while (0)

void set_last_error(const std::string& msg);
void log_error(const std::string& msg);  // "[NL] msg" on stderr + last error

// Runs f(); converts exceptions to an error code (never lets one escape).
template <typename F>
int guarded(const char* where, F&& f) {
    try {
        f();
        return NLHIP_OK;
    } catch (const Error& e) {
        log_error(std::string(where) + ": " + e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        log_error(std::string(where) + ": host allocation failed");
        return NLHIP_ERR_OOM;
    } catch (const std::exception& e) {
        log_error(std::string(where) + ": " + e.what());
        return NLHIP_ERR_INVALID;
    }
}

// ------------------------------------------------------------------ device

struct Device {
    int id = -1;
    int cu_count = 0;
    uint64_t hbm_bytes = 0;
    std::string name;
    std::string arch;
};

// Initialises the device on first call; throws Error(NLHIP_ERR_NO_DEVICE)
// when no gfx950 device is usable.
const Device& device();
bool device_usable();  // no throw

// Fault-injection hook for tests: the next `count` solves fail with
// NLHIP_ERR_INJECTED before touching the device.
bool take_injected_failure();

// ------------------------------------------------------------------ capture

// Stream capture and the runtime calls that may not overlap another thread's
// capture (allocation, stream/event creation, synchronous copies).  Graph
// capture holds this mutex exclusively from BeginCapture to EndCapture (it
// only records launches); those calls hold it shared.
std::shared_mutex& capture_mutex();
inline std::shared_lock<std::shared_mutex> no_capture_guard() {
    return std::shared_lock<std::shared_mutex>(capture_mutex());
}

// ------------------------------------------------------------------ buffers

template <typename T>
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        void* p = nullptr;
        hipError_t e;
        {
            auto g = no_capture_guard();
            e = hipMalloc(&p, n * sizeof(T));
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(NLHIP_ERR_OOM, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        p_ = static_cast<T*>(p);
        n_ = n;
    }
    void ensure(size_t n) { if (n_ < n || !p_) alloc(n); }
    void release() {
        if (p_) {
            auto g = no_capture_guard();
            (void)hipFree(p_);
        }
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

  private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

template <typename T>
class PinnedBuf {
  public:
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void ensure(size_t n) {
        if (n_ >= n && p_) return;
        release();
        if (n == 0) n = 1;
        void* p = nullptr;
        hipError_t e;
        {
            auto g = no_capture_guard();
            e = hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(NLHIP_ERR_OOM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
        }
        p_ = static_cast<T*>(p);
        n_ = n;
    }
    void release() {
        if (p_) {
            auto g = no_capture_guard();
            (void)hipHostFree(p_);
        }
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }

  private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// True when p is device (or managed) memory the GPU can address directly.
bool is_device_pointer(const void* p);

// Copies n elements into a device buffer from a host or device pointer.
void copy_in(void* dst, const void* src, size_t bytes, hipStream_t s);
// Copies n elements from device to a host or device pointer.
void copy_out(void* dst, const void* src, size_t bytes, hipStream_t s);

// ------------------------------------------------------------------ stream

class Stream {
  public:
    Stream();
    ~Stream();
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    hipStream_t get() const { return s_; }
    void sync() const { NLHIP_CHECK(hipStreamSynchronize(s_)); }

  private:
    hipStream_t s_ = nullptr;
};

class Event {
  public:
    Event();
    ~Event();
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    hipEvent_t get() const { return e_; }
    void record(hipStream_t s) { NLHIP_CHECK(hipEventRecord(e_, s)); }
    void sync() { NLHIP_CHECK(hipEventSynchronize(e_)); }

  private:
    hipEvent_t e_ = nullptr;
};

float elapsed_ms(const Event& a, const Event& b);

// ------------------------------------------------------------------ graphs

// A captured sequence of launches replayed as one hipGraph.
class Graph {
  public:
    Graph() = default;
    ~Graph() { reset(); }
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    template <typename F>
    void capture(hipStream_t s, F&& body) {
        reset();
        {
            std::unique_lock<std::shared_mutex> lock(capture_mutex());
            NLHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            NLHIP_CHECK(hipStreamEndCapture(s, &graph_));
        }
        NLHIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    }
    void launch(hipStream_t s) const { NLHIP_CHECK(hipGraphLaunch(exec_, s)); }
    bool ready() const { return exec_ != nullptr; }
    void reset() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
    }

  private:
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// Graphs of 4, 8, 16, ... iterations up to a cap K (launched as 4, 4, 8,
// ..., K, K, ...).  The host learns that a solve stopped one graph late, so
// a solve that stops after a few iterations replays few launches past its
// stopping test, and a long one still runs K iterations per graph.
//
// With NLHIP_DIRECT_LAUNCH=1 in the environment (read once), nothing is
// captured: launch() enqueues the same kernels one by one through the
// recording function of the latest ensure() call, which must come from the
// same scope as the launches (every caller does both in one function).
// Same kernels, same arguments, same results; for tools that trace kernel
// launches but not graph replays.
bool direct_launches();

class GraphLadder {
  public:
    void reset() {
        for (auto& g : g_) g.reset();
        n_ = 0;
        cap_ = 0;
        direct_ = nullptr;
    }
    // Captures the graphs for a cap of K on stream s (record(k) enqueues k
    // iterations); keeps them while K is unchanged (reset() when anything
    // baked into the launches changes).
    template <typename F>
    void ensure(uint32_t K, hipStream_t s, F&& record) {
        if (K == 0) K = 1;
        if (direct_launches()) {
            direct_ = std::function<void(uint32_t)>(record);
            if (n_ && K == cap_) return;
            n_ = 0;
            uint32_t k = K < kFirst ? K : kFirst;
            for (;;) {
                iters_[n_++] = k;
                if (k >= K || n_ == kLevels) break;
                k = 2 * k < K ? 2 * k : K;
            }
            cap_ = K;
            return;
        }
        if (n_ && K == cap_) return;
        reset();
        uint32_t k = K < kFirst ? K : kFirst;
        for (;;) {
            iters_[n_] = k;
            g_[n_].capture(s, [&] { record(k); });
            ++n_;
            if (k >= K || n_ == kLevels) break;
            k = 2 * k < K ? 2 * k : K;
        }
        cap_ = K;
    }
    bool ready() const { return n_ > 0; }
    uint32_t cap() const { return cap_; }
    // Iterations in the i-th launch (0-based) of a solve, and its launch.
    uint32_t iters(uint64_t i) const { return iters_[level(i)]; }
    void launch(uint64_t i, hipStream_t s) const {
        if (direct_) {
            direct_(iters(i));
            return;
        }
        g_[level(i)].launch(s);
    }

  private:
    static constexpr int kLevels = 12;
    static constexpr uint32_t kFirst = 4;
    int level(uint64_t i) const {
        return i == 0 ? 0 : int(i - 1 < uint64_t(n_ - 1) ? i - 1 : uint64_t(n_ - 1));
    }
    std::function<void(uint32_t)> direct_;
    Graph g_[kLevels];
    uint32_t iters_[kLevels] = {};
    int n_ = 0;
    uint32_t cap_ = 0;
};

// ------------------------------------------------------------------ system

// What an SpMV launch reads: row pointer, column indices, values, and the
// resolved engine code (NLHIP_RPW_SWITCH).
template <typename T>
struct SpmvOps {
    const uint32_t* rp;
    Cols cols;
    const T* val;
    int code;
};

template <typename T>
struct CGWorkspace {
    uint32_t n = 0, P = 0;
    int code = 0;  // SpMV engine code the arguments were set up for
    DevBuf<T> x, b, r, h, d, q, part;
    DevBuf<CGStatus<T>> st;
    PinnedBuf<CGStatus<T>> st_host;  // two slots
    GraphLadder graph;
    int graph_rpw = 0;
    bool graph_precond = false;
    CGArgs<T> graph_args{};  // arguments baked into the captured graph
    CGArgs<T> args{};
    void ensure(uint32_t n_, uint32_t P_);
};

// Workspace of a multi right-hand-side solve (k_spmm.hip), n x k interleaved.
template <typename T>
struct MultiWork {
    uint32_t n = 0, k = 0, P = 0;
    DevBuf<T> x, b, r, h, d, q, part, col;  // part: 4 x P x k; col: 6 x k
    DevBuf<int32_t> iters, done, skip, again;
    DevBuf<int32_t> xnz;                    // X has a nonzero (args.x_nonzero)
    DevBuf<T> chk;                          // 2 x k: |b - A x|^2, |b|^2 (confirm)
    PinnedBuf<int32_t> host;                // two polling slots of 32 done flags
    PinnedBuf<int32_t> iters_host;          // k iteration counts (refinement)
    GraphLadder graph;
    MultiArgs<T> args{};
    MultiArgs<T> graph_args{};
    int graph_max_iter = -1;
    bool graph_precond = false;
};

// Mixed-precision refinement (single, multi-RHS, batched): it ends, back at
// the iterate with the lowest FP64 residual, after kRefineNoGain steps in a
// row that do not lower that residual, or at once when the squared residual
// exceeds kRefineDiverged times the lowest (or is NaN / Inf).
constexpr uint32_t kRefineNoGain = 3;
constexpr double kRefineDiverged = 1e4;

// An FP64 PCG solve on a system's ws64 that can stop and continue (slices of
// nlhipSolveAdvance, the confirmation of the true residual).  `cur` is the
// host copy of the device status after the last stop; the device continues
// from it bitwise.
struct PcgState {
    bool active = false;    // a solve is open on the system's ws64
    bool ran = false;       // the loop has been entered at least once
    bool finished = false;  // stopped for good (test met and confirmed, cap)
    bool precond = false;
    nlhip_solve_params p{};
    uint32_t K = 0;         // iterations per captured graph
    CGStatus<double> cur{};  // device status of the current recurrence
    uint32_t base = 0;      // iterations of the recurrences before it (restarts)
    double thresh = 0;      // recursive threshold in force
    uint32_t first_stop = ~0u;  // iterations when the recursive test first held
    double rr = -1, bb = 0;     // true |b - A x|^2 and |b|^2 at the last check
    double last_rr = -1;
    int rounds = 0;
    int last_kind = 0;      // last confirmation round: 1 continued, 2 restarted
    uint32_t restarts = 0;
    double device_seconds = 0;
    uint32_t total() const { return base + uint32_t(cur.iters); }
};

// The device buffers of one FP64/FP32 PCG: iterate, right-hand side, the
// recurrence's vectors, per-workgroup partials ((kSpmvGridMul + 3) x grid)
// and the status block.
template <typename T>
struct CGBuffers {
    T *x, *b, *r, *h, *d, *q, *part;
    CGStatus<T>* st;
};

// One way of running the device loop of an FP64 PCG (captured graphs, or
// direct launches without any allocation); pcg_drive holds the stopping and
// confirmation logic common to them.
struct PcgOps {
    virtual ~PcgOps() = default;
    // initial residual from the current x, status read back
    virtual CGStatus<double> start(double eps) = 0;
    // stop/done cleared, threshold and cap set (continues bitwise)
    virtual void reopen(const CGStatus<double>& cur, double thresh, int cap) = 0;
    // iterations until the device loop stops, K per poll, at most ~budget
    virtual CGStatus<double> run(uint32_t K, uint64_t budget) = 0;
    // |b - A x|^2 and |b|^2 in FP64
    virtual void norms(double& rr, double& bb) = 0;
};

// A CRS matrix resident on the device, with its Jacobi inverse diagonal and
// (optionally) FP32 copies, its own stream and reusable solve workspaces.
struct System {
    uint32_t n = 0;
    uint64_t nnz = 0;
    DevBuf<uint32_t> rowptr, colind;
    DevBuf<uint16_t> ci16;   // 16-bit column offsets (empty when the span is too wide)
    DevBuf<uint32_t> segb;   // their per-segment bases
    // 2x2 block form (BlockStreamEngine; empty when the matrix has none):
    // block row pointer, block columns as 16-bit segment offsets (bci16,
    // bsegb) or, when a segment spans more than 16 bits, as 32-bit indices
    // (bcol), values
    DevBuf<uint32_t> brp, bsegb, bcol;
    DevBuf<uint16_t> bci16;
    DevBuf<double> bval;
    DevBuf<float> bval32;
    DevBuf<double> val, dinv;
    DevBuf<float> val32, dinv32;
    int rpw = 64;        // rows per wavefront of the SpMV engine
    uint32_t grid = 1;   // workgroups per launch
    Stream stream;
    std::mutex mu;
    CGWorkspace<double> ws64;
    CGWorkspace<float> ws32;
    DevBuf<double> scratch;  // residual vector + partials for refinement
    MultiWork<double> m64;   // multi right-hand-side workspaces
    MultiWork<float> m32;
    DevBuf<double> mscale;
    PcgState pcg;            // the FP64 solve open on ws64 (nlhipSolveBegin)
    PinnedBuf<double> host_scalars;

    void ensure_fp32();
    uint64_t device_bytes() const;
    Cols cols() const { return Cols{colind.get(), ci16.get(), segb.get()}; }
    bool blocks() const { return bval.get() != nullptr; }
    // Arrays and engine code of the SpMV for the current launch shape (FP64
    // or, for T = float, the FP32 values; ensure_fp32() first).
    template <typename T>
    SpmvOps<T> ops() const {
        const T* v;
        const T* bv;
        if constexpr (std::is_same<T, double>::value) {
            v = val.get();
            bv = bval.get();
        } else {
            v = val32.get();
            bv = bval32.get();
        }
        if (blocks() && (rpw >= 73 || (rpw == 64 && n >= kBlockMinRows))) {
            if (bci16.get())
                return SpmvOps<T>{brp.get(), Cols{nullptr, bci16.get(), bsegb.get()}, bv,
                                  block_engine(rpw >= 73 ? rpw : 73, true)};
            return SpmvOps<T>{brp.get(), Cols{bcol.get(), nullptr, nullptr}, bv,
                              block_engine(rpw >= 73 ? rpw : 73, false)};
        }
        const int code = rpw >= 73 ? 64 : rpw;
        return SpmvOps<T>{rowptr.get(), cols(), v, engine_rpw(code, cols())};
    }
};

// Builds the 2x2 block form (brp, bval, and the block columns as 16-bit
// offsets bci16 / bsegb or, when a segment's span does not fit, as 32-bit
// indices bcol; see BlockStreamEngine) of a device CRS matrix when it has
// that form; false otherwise.
bool build_block_form(uint32_t n, uint64_t nnz, const uint32_t* rowptr, const uint32_t* colind,
                      const double* val, hipStream_t s, DevBuf<uint32_t>& brp,
                      DevBuf<uint16_t>& bci16, DevBuf<uint32_t>& bsegb, DevBuf<uint32_t>& bcol,
                      DevBuf<double>& bval);
bool build_blocks(System& S);

// Builds the 16-bit column offsets of a device colind (see Cols); false when
// a segment spans more than 65535 columns.
bool build_col_offsets(uint64_t nnz, const uint32_t* colind, hipStream_t s,
                       DevBuf<uint16_t>& ci16, DevBuf<uint32_t>& segb);

// Launch shape of the eigensolver's multi-vector SpMM (k_eigen.hip).
struct EigShape {
    uint32_t grid = 1, rows_per_block = 64;
    bool one_chunk = false;  // spmm_one_chunk for the block width
};

// Launch-shape autotuning of a system's CG (S.mu held).
struct TuneResult {
    int rpw;
    uint32_t grid;
    double seconds;  // per CG iteration
};
void profile_cg(System& S, bool fp32, uint32_t iters, double t[3]);
TuneResult autotune(System& S, bool fp32, uint32_t iters);

// Validates and uploads; throws Error.
System* system_create(uint32_t n, const uint32_t* rowptr, const uint32_t* colind,
                      const double* val, uint32_t flags);
// Takes over a device CRS built by this library (already valid: sorted
// columns in range, finite values) and finishes the system as
// system_create does.  Work on `producer` must be complete.
System* system_adopt(uint32_t n, uint64_t nnz, DevBuf<uint32_t>&& rowptr,
                     DevBuf<uint32_t>&& colind, DevBuf<double>&& val, uint32_t flags);

// Least-squares normal equations on the device (k_assemble.hip).  Device
// arrays: rows rp[nrows + 1] over entries (var, coef)[nent]; row right-hand
// sides row_rhs[nsys][nrows] (null: 0); free_index[nvars] (0xffffffff for a
// locked variable); values[nsys][nvars] of the locked variables (null: 0).
struct LSInput {
    uint32_t nrows = 0, nvars = 0, nsys = 1, nfree = 0;
    uint64_t nent = 0;
    const uint32_t* rp = nullptr;
    const uint32_t* var = nullptr;
    const double* coef = nullptr;
    const double* row_rhs = nullptr;
    const uint32_t* free_index = nullptr;
    const double* values = nullptr;
};
struct LSOutput {
    DevBuf<uint32_t> rowptr, colind;  // CRS over the free variables, sorted columns
    DevBuf<double> val;
    DevBuf<double> rhs;               // nsys x nfree
    uint64_t nnz = 0;
};
void ls_assemble_device(const LSInput& in, LSOutput& out, hipStream_t s);
// LSCM rows of a triangle mesh on the device (points nv x 3, triangles nt x 3,
// device arrays): two rows of 6 entries per non-degenerate triangle, in
// triangle order, over the variables (2 v, 2 v + 1) = (u_v, v_v).
struct LSRows {
    uint32_t nrows = 0;
    DevBuf<uint32_t> rp, var;
    DevBuf<double> coef;
};
void lscm_rows_device(uint32_t nv, const double* points, uint32_t nt, const uint32_t* triangles,
                      LSRows& rows, hipStream_t s);
// Throws when a triangle's vertices lie in different charts (device arrays;
// chart c owns vertices [voff[c], voff[c+1])).
void check_triangle_charts(uint32_t nt, const uint32_t* T, const uint32_t* voff, uint32_t nc,
                           hipStream_t s);
// nlhipLeastSquaresCreate: host or device inputs, checks, assembly, system;
// rhs (nsys x nfree, host or device) may be null.
System* ls_create(uint32_t nvars, uint32_t nrows, const uint32_t* row_ptr, const uint32_t* var,
                  const double* coef, uint32_t nsys, const double* row_rhs,
                  const uint8_t* locked, const double* values, uint32_t flags, double* rhs,
                  uint32_t* nfree_out);
void default_launch(uint32_t n, uint64_t nnz, int& rpw, uint32_t& grid);

// |b - A x| / |b| in FP64 on the device.
double true_rel_residual(System& S, const double* b_dev, const double* x_dev);
// CG launch arguments of system S over the buffers w (code: SpMV engine).
template <typename T>
void fill_cg_args(const System& S, const CGBuffers<T>& w, const T* dinv, CGArgs<T>& a,
                  int& code);
// Iterations per status poll of a CG on S (about 2 ms of device work).
uint32_t auto_check_every(const System& S);
// The FP64 PCG driver: open (initial residual), run to the end or to
// `until` iterations in all, statistics.
void pcg_open(PcgState& P, PcgOps& ops, const nlhip_solve_params& p, bool precond, uint32_t K);
void pcg_drive(PcgState& P, PcgOps& ops, uint64_t until);
void pcg_stats(const PcgState& P, nlhip_solve_stats& st);
// Full solve with staging of host/device b and x (takes S.mu).  dinv: the
// Jacobi inverse diagonal of a separate preconditioner object
// (nlSolveSystemIterative's P) instead of the system's own; null: the
// system's.
void solve(System& S, const double* b, double* x, const nlhip_solve_params& p,
           nlhip_solve_stats& st, const double* dinv = nullptr);

// k right-hand sides at once (B, X column-major n x k, host or device).
// per (k entries) and total may be null.
void solve_multi(System& S, uint32_t k, const double* B, double* X,
                 const nlhip_solve_params& p, nlhip_solve_stats* per, nlhip_solve_stats* total);

// True when every pass of the SpMM row engine with G lanes per row (G =
// columns rounded up to a power of two) fits one LDS chunk on the matrix
// with row pointer rp (device), so the engine can finish rows one by one.
bool spmm_one_chunk(uint32_t m, const uint32_t* rp, uint32_t G, hipStream_t s);

// ------------------------------------------------------------------ SSOR

// Arguments of the SSOR sweeps (k_ssor.hip).  Each sweep lists its rows by
// level (positions), and for each position the entries that sweep reads
// (j < i forward, j > i backward) packed in CRS order.
struct SsorSweepArgs {
    const uint32_t* rows;  // row at each position
    const uint32_t* off;   // first position of each level (levels + 1)
    const uint32_t* eoff;  // first packed entry of each position (n + 1)
    const uint32_t* col;
    const double* val;
};
struct SsorArgs {
    SsorSweepArgs lo, up;
    const double* diag;    // d_i
    const double* dscale;  // d_i / omega
    const double* x;       // input
    double* w;             // forward sweep
    double* y;             // output
    double omega, post;    // post = 2 - omega
    uint32_t n;
};
hipError_t ssor_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* val,
                     double omega, double* diag, double* dscale, uint32_t* bad, hipStream_t s);
// Levels [l0, l1) of one sweep; `rows` rows in all (one level: the grid
// covers them; several levels: one workgroup).
hipError_t ssor_sweep(const SsorArgs& a, bool upper, uint32_t l0, uint32_t l1, uint32_t rows,
                      hipStream_t s);
hipError_t ssor_scale(const SsorArgs& a, hipStream_t s);

// An SSOR(omega) preconditioner of a resident matrix (its own copy of the
// CRS, level schedules built on the host, the sweeps captured as a graph).
class Ssor;
Ssor* ssor_create(System& S, double omega);
void ssor_destroy(Ssor* P);
uint32_t ssor_n(const Ssor& P);
// y = SSOR^-1 x for device vectors, on stream s.
void ssor_apply(Ssor& P, const double* x, double* y, hipStream_t s);

// The preconditioner of the host-driven solvers: Jacobi (dinv), SSOR, or
// none.
struct Precond {
    const double* dinv = nullptr;
    Ssor* ssor = nullptr;
};
void precond_apply(const Precond& P, uint32_t n, const double* in, double* out, hipStream_t s);

// Device BiCGSTAB on BLAS-1 kernels (NL_BICGSTAB); returns iterations.
uint32_t bicgstab_device(System& S, const Precond& P, const double* b_dev, double* x_dev,
                         double eps, uint32_t max_iter, double* rel2);
// Restarted GMRES(m) on the device (right-preconditioned when P is set);
// returns the number of Arnoldi steps.
uint32_t gmres_device(System& S, const Precond& P, const double* b_dev, double* x_dev,
                      double eps, uint32_t max_iter, uint32_t m, double* rel2);
// PCG with any Precond, host-driven like BiCGSTAB (the fused CG kernels
// apply Jacobi only); returns iterations.
uint32_t pcg_device(System& S, const Precond& P, const double* b_dev, double* x_dev, double eps,
                    uint32_t max_iter, double* rel2);

}  // namespace nlhip
