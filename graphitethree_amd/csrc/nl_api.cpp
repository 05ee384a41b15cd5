// OpenNL C API on top of the HIP runtime: matrices, NLBlas tables,
// nlSolveSystemIterative, the HIP extension entry points and the context
// (least-squares) API.  Declared in include/nl.h and include/nl_hip.h.
//
// Host-side work here is limited to what OpenNL does on the host before a
// solve: collecting coefficients, forming least-squares normal equations and
// compressing rows to CRS.  Every matrix-vector product, BLAS-1 operation,
// preconditioner application and solve runs on the GPU; when no gfx950
// device is usable they report an error instead of computing on the CPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "nlhip_internal.h"

using namespace nlhip;

namespace {

constexpr uint32_t kMagicHIP = 0x48495043u;     // device CRS matrix
constexpr uint32_t kMagicJacobi = 0x4a41434fu;  // Jacobi preconditioner
constexpr uint32_t kMagicSSOR = 0x53534f52u;    // SSOR preconditioner

// Device mirror of a host matrix (dynamic or CRS); rebuilt when stale.
struct Mirror {
    System* sys = nullptr;
    NLuint version = 0;
    ~Mirror() { delete sys; }
};

struct HIPMatrix {
    NLuint m, n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
    uint32_t magic;
    System* sys;
};

struct JacobiMatrix {
    NLuint m, n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
    uint32_t magic;
    DevBuf<double>* dinv;
};

struct SSORMatrix {
    NLuint m, n;
    NLenum type;
    NLDestroyMatrixFunc destroy_func;
    NLMultMatrixVectorFunc mult_func;
    uint32_t magic;
    Ssor* ssor;
};

// Utility stream and staging buffers for operations not tied to a system
// (host-vector BLAS, preconditioner application).
struct Util {
    Stream stream;
    std::mutex mu;
    DevBuf<double> a, b, part, out;
    PinnedBuf<double> host;
    void ensure(size_t n) {
        a.ensure(n);
        b.ensure(n);
        part.ensure(4096);
        out.ensure(2);
        host.ensure(2);
    }
};

Util& util() {
    static Util* u = new Util();  // intentionally never destroyed (process lifetime)
    return *u;
}

// -------------------------------------------------------------- dynamic

void dyn_destroy(NLMatrix M);
void dyn_mult(NLMatrix M, const double* x, double* y);
void crs_destroy(NLMatrix M);
void crs_mult(NLMatrix M, const double* x, double* y);
void hip_destroy(NLMatrix M);
void hip_mult(NLMatrix M, const double* x, double* y);
void jac_destroy(NLMatrix M);
void jac_mult(NLMatrix M, const double* x, double* y);

void row_add(NLRowColumn* R, NLuint j, double a) {
    for (NLuint k = 0; k < R->size; ++k) {
        if (R->coeff[k].index == j) {
            R->coeff[k].value += a;
            return;
        }
    }
    if (R->size == R->capacity) {
        NLuint cap = R->capacity ? 2 * R->capacity : 4;
        NLCoeff* c = static_cast<NLCoeff*>(std::realloc(R->coeff, sizeof(NLCoeff) * cap));
        if (!c) throw std::bad_alloc();
        R->coeff = c;
        R->capacity = cap;
    }
    R->coeff[R->size].index = j;
    R->coeff[R->size].value = a;
    ++R->size;
}

NLSparseMatrix* dyn_new(NLuint m, NLuint n) {
    NLSparseMatrix* M = static_cast<NLSparseMatrix*>(std::calloc(1, sizeof(NLSparseMatrix)));
    if (!M) throw std::bad_alloc();
    M->m = m;
    M->n = n;
    M->type = NL_MATRIX_SPARSE_DYNAMIC;
    M->destroy_func = dyn_destroy;
    M->mult_func = dyn_mult;
    M->storage = NL_MATRIX_STORE_ROWS;
    M->diag_size = std::min(m, n);
    M->row = static_cast<NLRowColumn*>(std::calloc(m ? m : 1, sizeof(NLRowColumn)));
    if (!M->row) {
        std::free(M);
        throw std::bad_alloc();
    }
    return M;
}

void dyn_destroy(NLMatrix Mg) {
    NLSparseMatrix* M = reinterpret_cast<NLSparseMatrix*>(Mg);
    for (NLuint i = 0; i < M->m; ++i) std::free(M->row[i].coeff);
    std::free(M->row);
    std::free(M->diag);
    delete static_cast<Mirror*>(M->hip_mirror);
    M->row = nullptr;
    M->hip_mirror = nullptr;
}

// Rows compressed to CRS arrays with increasing columns.
void dyn_to_arrays(const NLSparseMatrix* M, std::vector<NLuint>& rp, std::vector<NLuint>& ci,
                   std::vector<double>& va) {
    rp.assign(size_t(M->m) + 1, 0);
    uint64_t nnz = 0;
    for (NLuint i = 0; i < M->m; ++i) {
        nnz += M->row[i].size;
        if (nnz > 0xFFFFFFFFull) throw Error(NLHIP_ERR_UNSUPPORTED, "more than 2^32-1 nonzeros");
        rp[i + 1] = NLuint(nnz);
    }
    ci.resize(nnz);
    va.resize(nnz);
    std::vector<NLCoeff> tmp;
    for (NLuint i = 0; i < M->m; ++i) {
        const NLRowColumn& R = M->row[i];
        tmp.assign(R.coeff, R.coeff + R.size);
        std::sort(tmp.begin(), tmp.end(),
                  [](const NLCoeff& a, const NLCoeff& b) { return a.index < b.index; });
        for (NLuint k = 0; k < R.size; ++k) {
            ci[rp[i] + k] = tmp[k].index;
            va[rp[i] + k] = tmp[k].value;
        }
    }
}

NLCRSMatrix* crs_from_arrays(NLuint m, NLuint n, const NLuint* rp, const NLuint* ci,
                             const double* va) {
    NLCRSMatrix* C = static_cast<NLCRSMatrix*>(std::calloc(1, sizeof(NLCRSMatrix)));
    if (!C) throw std::bad_alloc();
    const uint64_t nnz = rp[m];
    C->m = m;
    C->n = n;
    C->type = NL_MATRIX_CRS;
    C->destroy_func = crs_destroy;
    C->mult_func = crs_mult;
    C->rowptr = static_cast<NLuint*>(std::malloc(sizeof(NLuint) * (size_t(m) + 1)));
    C->colind = static_cast<NLuint*>(std::malloc(sizeof(NLuint) * (nnz ? nnz : 1)));
    C->val = static_cast<double*>(std::malloc(sizeof(double) * (nnz ? nnz : 1)));
    if (!C->rowptr || !C->colind || !C->val) throw std::bad_alloc();
    std::memcpy(C->rowptr, rp, sizeof(NLuint) * (size_t(m) + 1));
    if (nnz) {
        std::memcpy(C->colind, ci, sizeof(NLuint) * nnz);
        std::memcpy(C->val, va, sizeof(double) * nnz);
    }
    C->nslices = 1;
    return C;
}

void crs_destroy(NLMatrix Mg) {
    NLCRSMatrix* C = reinterpret_cast<NLCRSMatrix*>(Mg);
    std::free(C->rowptr);
    std::free(C->colind);
    std::free(C->val);
    std::free(C->sliceptr);
    delete static_cast<System*>(C->hip_mirror);
    C->rowptr = C->colind = nullptr;
    C->val = nullptr;
    C->hip_mirror = nullptr;
}

// The device system behind any solvable matrix (uploads on first use).
// Serialises the creation and rebuild of host matrices' device mirrors: two
// threads multiplying or solving with one matrix (read-only use, which
// OpenNL allows) must not both build its mirror, nor use one that another
// thread is replacing.  A matrix that is modified while another thread uses
// it is outside what OpenNL allows, as upstream.
static std::mutex& mirror_mutex() {
    static std::mutex* m = new std::mutex();  // process lifetime
    return *m;
}

System* system_of(NLMatrix M) {
    if (!M) throw Error(NLHIP_ERR_INVALID, "null matrix");
    if (M->type == NL_MATRIX_CRS) {
        NLCRSMatrix* C = reinterpret_cast<NLCRSMatrix*>(M);
        if (C->m != C->n) throw Error(NLHIP_ERR_UNSUPPORTED, "non-square CRS matrix");
        std::lock_guard<std::mutex> lock(mirror_mutex());
        if (!C->hip_mirror) C->hip_mirror = system_create(C->m, C->rowptr, C->colind, C->val, 0);
        return static_cast<System*>(C->hip_mirror);
    }
    if (M->type == NL_MATRIX_SPARSE_DYNAMIC) {
        NLSparseMatrix* D = reinterpret_cast<NLSparseMatrix*>(M);
        if (D->m != D->n) throw Error(NLHIP_ERR_UNSUPPORTED, "non-square dynamic matrix");
        std::lock_guard<std::mutex> lock(mirror_mutex());
        Mirror* mi = static_cast<Mirror*>(D->hip_mirror);
        if (!mi) {
            mi = new Mirror();
            D->hip_mirror = mi;
        }
        if (!mi->sys || mi->version != D->version) {
            std::vector<NLuint> rp, ci;
            std::vector<double> va;
            dyn_to_arrays(D, rp, ci, va);
            delete mi->sys;
            mi->sys = nullptr;
            mi->sys = system_create(D->m, rp.data(), ci.data(), va.data(), 0);
            mi->version = D->version;
        }
        return mi->sys;
    }
    if (M->type == NL_MATRIX_OTHER) {
        HIPMatrix* H = reinterpret_cast<HIPMatrix*>(M);
        if (H->magic == kMagicHIP) return H->sys;
    }
    throw Error(NLHIP_ERR_UNSUPPORTED, "matrix type has no GPU implementation");
}

SSORMatrix* as_ssor(NLMatrix P) {
    if (P && P->type == NL_MATRIX_OTHER) {
        SSORMatrix* Q = reinterpret_cast<SSORMatrix*>(P);
        if (Q->magic == kMagicSSOR) return Q;
    }
    return nullptr;
}

JacobiMatrix* as_jacobi(NLMatrix P) {
    if (P && P->type == NL_MATRIX_OTHER) {
        JacobiMatrix* J = reinterpret_cast<JacobiMatrix*>(P);
        if (J->magic == kMagicJacobi) return J;
    }
    return nullptr;
}

// The preconditioner object P of nlSolveSystemIterative for a system of
// size n: Jacobi or SSOR (or none when P is null).
Precond precond_of(NLMatrix P, uint32_t n) {
    Precond pc;
    if (!P) return pc;
    if (SSORMatrix* Q = as_ssor(P)) {
        if (Q->n != n) throw Error(NLHIP_ERR_INVALID, "preconditioner size mismatch");
        pc.ssor = Q->ssor;
        return pc;
    }
    JacobiMatrix* J = as_jacobi(P);
    if (!J) throw Error(NLHIP_ERR_UNSUPPORTED, "only Jacobi and SSOR preconditioners run on the GPU");
    if (J->n != n) throw Error(NLHIP_ERR_INVALID, "preconditioner size mismatch");
    pc.dinv = J->dinv->get();
    return pc;
}

// y = A x for a device system; x, y host or device.
void system_mult(System* S, const double* x, double* y) {
    std::lock_guard<std::mutex> lock(S->mu);
    hipStream_t s = S->stream.get();
    S->ws64.ensure(S->n, S->grid);
    copy_in(S->ws64.x.get(), x, size_t(S->n) * 8, s);
    const SpmvOps<double> o = S->ops<double>();
    NLHIP_CHECK(spmv<double>(S->n, o.rp, o.cols, o.val, S->ws64.x.get(), S->ws64.q.get(), o.code,
                             S->grid, s));
    copy_out(y, S->ws64.q.get(), size_t(S->n) * 8, s);
    S->stream.sync();
}

void dyn_mult(NLMatrix M, const double* x, double* y) {
    guarded("nlMultMatrixVector", [&] { system_mult(system_of(M), x, y); });
}
void crs_mult(NLMatrix M, const double* x, double* y) {
    guarded("nlMultMatrixVector", [&] { system_mult(system_of(M), x, y); });
}
void hip_mult(NLMatrix M, const double* x, double* y) {
    guarded("nlMultMatrixVector", [&] { system_mult(system_of(M), x, y); });
}

void hip_destroy(NLMatrix M) {
    HIPMatrix* H = reinterpret_cast<HIPMatrix*>(M);
    delete H->sys;
    H->sys = nullptr;
}

void jac_destroy(NLMatrix M) {
    JacobiMatrix* J = reinterpret_cast<JacobiMatrix*>(M);
    delete J->dinv;
    J->dinv = nullptr;
}

void jac_mult(NLMatrix M, const double* x, double* y) {
    guarded("nlMultMatrixVector", [&] {
        JacobiMatrix* J = reinterpret_cast<JacobiMatrix*>(M);
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        hipStream_t s = u.stream.get();
        u.ensure(J->n);
        copy_in(u.a.get(), x, size_t(J->n) * 8, s);
        NLHIP_CHECK(blas_dmul(J->n, J->dinv->get(), u.a.get(), u.b.get(), s));
        copy_out(y, u.b.get(), size_t(J->n) * 8, s);
        u.stream.sync();
    });
}

void ssor_matrix_destroy(NLMatrix M) {
    SSORMatrix* Q = reinterpret_cast<SSORMatrix*>(M);
    ssor_destroy(Q->ssor);
    Q->ssor = nullptr;
}

void ssor_matrix_mult(NLMatrix M, const double* x, double* y) {
    guarded("nlMultMatrixVector", [&] {
        SSORMatrix* Q = reinterpret_cast<SSORMatrix*>(M);
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        hipStream_t s = u.stream.get();
        u.ensure(Q->n);
        copy_in(u.a.get(), x, size_t(Q->n) * 8, s);
        ssor_apply(*Q->ssor, u.a.get(), u.b.get(), s);
        copy_out(y, u.b.get(), size_t(Q->n) * 8, s);
        u.stream.sync();
    });
}

NLMatrix ssor_new(NLMatrix M, double omega) {
    System* S = system_of(M);
    auto Q = std::make_unique<SSORMatrix>();
    Q->m = Q->n = S->n;
    Q->type = NL_MATRIX_OTHER;
    Q->destroy_func = ssor_matrix_destroy;
    Q->mult_func = ssor_matrix_mult;
    Q->magic = kMagicSSOR;
    Q->ssor = ssor_create(*S, omega);
    return reinterpret_cast<NLMatrix>(Q.release());
}

NLMatrix jacobi_new(NLMatrix M) {
    System* S = system_of(M);
    auto J = std::make_unique<JacobiMatrix>();
    J->m = J->n = S->n;
    J->type = NL_MATRIX_OTHER;
    J->destroy_func = jac_destroy;
    J->mult_func = jac_mult;
    J->magic = kMagicJacobi;
    J->dinv = new DevBuf<double>(S->n);
    NLHIP_CHECK(hipMemcpyAsync(J->dinv->get(), S->dinv.get(), size_t(S->n) * 8,
                               hipMemcpyDeviceToDevice, S->stream.get()));
    S->stream.sync();
    return reinterpret_cast<NLMatrix>(J.release());
}

// ------------------------------------------------------------------ BLAS

struct BlasImpl {
    NLBlas table;
    bool device_vectors;
};

void check_inc(int incx, int incy = 1) {
    if (incx != 1 || incy != 1) throw Error(NLHIP_ERR_UNSUPPORTED, "BLAS stride other than 1");
}

void* b_malloc(NLBlas_t blas, NLmemoryType type, size_t size) {
    void* p = nullptr;
    guarded("NLBlas::Malloc", [&] {
        if (type == NL_DEVICE_MEMORY) {
            device();
            auto g = no_capture_guard();
            NLHIP_CHECK(hipMalloc(&p, size ? size : 1));
        } else {
            p = std::malloc(size ? size : 1);
            if (!p) throw std::bad_alloc();
        }
        blas->used_ram[type] += size;
        blas->max_used_ram[type] = std::max(blas->max_used_ram[type], blas->used_ram[type]);
    });
    return p;
}

void b_free(NLBlas_t blas, NLmemoryType type, size_t size, void* p) {
    if (!p) return;
    if (type == NL_DEVICE_MEMORY) {
        auto g = no_capture_guard();
        (void)hipFree(p);
    } else {
        std::free(p);
    }
    blas->used_ram[type] -= std::min<NLulong>(blas->used_ram[type], size);
}

void b_memcpy(NLBlas_t, void* to, NLmemoryType, const void* from, NLmemoryType, size_t size) {
    guarded("NLBlas::Memcpy", [&] {
        device();
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        NLHIP_CHECK(hipMemcpyAsync(to, from, size, hipMemcpyDefault, u.stream.get()));
        u.stream.sync();
    });
}

// Host-vector BLAS: vectors are staged to the device and back.
void h_dcopy(NLBlas_t, int n, const double* x, int incx, double* y, int incy) {
    guarded("nlHostBlas()->Dcopy", [&] {
        check_inc(incx, incy);
        device();
        if (n > 0) std::memmove(y, x, sizeof(double) * size_t(n));
    });
}

void h_dscal(NLBlas_t blas, int n, double a, double* x, int incx) {
    guarded("nlHostBlas()->Dscal", [&] {
        check_inc(incx);
        if (n <= 0) return;
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        hipStream_t s = u.stream.get();
        u.ensure(size_t(n));
        copy_in(u.a.get(), x, size_t(n) * 8, s);
        NLHIP_CHECK(blas_dscal(size_t(n), a, u.a.get(), s));
        copy_out(x, u.a.get(), size_t(n) * 8, s);
        u.stream.sync();
        blas->flops += NLulong(n);
    });
}

void h_daxpy(NLBlas_t blas, int n, double a, const double* x, int incx, double* y, int incy) {
    guarded("nlHostBlas()->Daxpy", [&] {
        check_inc(incx, incy);
        if (n <= 0) return;
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        hipStream_t s = u.stream.get();
        u.ensure(size_t(n));
        copy_in(u.a.get(), x, size_t(n) * 8, s);
        copy_in(u.b.get(), y, size_t(n) * 8, s);
        NLHIP_CHECK(blas_daxpy(size_t(n), a, u.a.get(), u.b.get(), s));
        copy_out(y, u.b.get(), size_t(n) * 8, s);
        u.stream.sync();
        blas->flops += 2 * NLulong(n);
    });
}

double dot_staged(int n, const double* x, const double* y, bool sqrt_result) {
    Util& u = util();
    std::lock_guard<std::mutex> lock(u.mu);
    hipStream_t s = u.stream.get();
    u.ensure(size_t(n));
    const double* dx = x;
    const double* dy = y;
    if (!is_device_pointer(x)) {
        copy_in(u.a.get(), x, size_t(n) * 8, s);
        dx = u.a.get();
    }
    if (y == x) {
        dy = dx;
    } else if (!is_device_pointer(y)) {
        copy_in(u.b.get(), y, size_t(n) * 8, s);
        dy = u.b.get();
    }
    uint64_t g = (uint64_t(n) + 255) / 256;
    const uint32_t grid = uint32_t(std::clamp<uint64_t>(g, 1, 1024));
    NLHIP_CHECK(blas_ddot(size_t(n), dx, dy, u.part.get(), grid, u.out.get(), s));
    NLHIP_CHECK(hipMemcpyAsync(u.host.get(), u.out.get(), 2 * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    u.stream.sync();
    return sqrt_result ? u.host.get()[1] : u.host.get()[0];
}

double h_ddot(NLBlas_t blas, int n, const double* x, int incx, const double* y, int incy) {
    double r = 0.0;
    guarded("NLBlas::Ddot", [&] {
        check_inc(incx, incy);
        if (n <= 0) return;
        r = dot_staged(n, x, y, false);
        blas->flops += 2 * NLulong(n);
    });
    return r;
}

double h_dnrm2(NLBlas_t blas, int n, const double* x, int incx) {
    double r = 0.0;
    guarded("NLBlas::Dnrm2", [&] {
        check_inc(incx);
        if (n <= 0) return;
        r = dot_staged(n, x, x, true);
        blas->flops += 2 * NLulong(n);
    });
    return r;
}

void h_dmul(NLBlas_t blas, int n, const double* a, const double* x, double* y) {
    guarded("nlHostBlas()->Dmul", [&] {
        if (n <= 0) return;
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        hipStream_t s = u.stream.get();
        u.ensure(size_t(n));
        copy_in(u.a.get(), a, size_t(n) * 8, s);
        copy_in(u.b.get(), x, size_t(n) * 8, s);
        NLHIP_CHECK(blas_dmul(size_t(n), u.a.get(), u.b.get(), u.b.get(), s));
        copy_out(y, u.b.get(), size_t(n) * 8, s);
        u.stream.sync();
        blas->flops += NLulong(n);
    });
}

void h_dzero(NLBlas_t, int n, double* x) {
    guarded("nlHostBlas()->Dzero", [&] {
        device();
        if (n > 0) std::memset(x, 0, sizeof(double) * size_t(n));
    });
}

// Device-vector BLAS (nlHIPBlas): pointers are device memory.
void d_dcopy(NLBlas_t, int n, const double* x, int incx, double* y, int incy) {
    guarded("nlHIPBlas()->Dcopy", [&] {
        check_inc(incx, incy);
        if (n > 0) {
            Util& u = util();
            std::lock_guard<std::mutex> lock(u.mu);
            NLHIP_CHECK(hipMemcpyAsync(y, x, sizeof(double) * size_t(n), hipMemcpyDefault,
                                       u.stream.get()));
            u.stream.sync();
        }
    });
}

void d_dscal(NLBlas_t blas, int n, double a, double* x, int incx) {
    guarded("nlHIPBlas()->Dscal", [&] {
        check_inc(incx);
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        NLHIP_CHECK(blas_dscal(size_t(std::max(n, 0)), a, x, u.stream.get()));
        u.stream.sync();
        blas->flops += NLulong(std::max(n, 0));
    });
}

void d_daxpy(NLBlas_t blas, int n, double a, const double* x, int incx, double* y, int incy) {
    guarded("nlHIPBlas()->Daxpy", [&] {
        check_inc(incx, incy);
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        NLHIP_CHECK(blas_daxpy(size_t(std::max(n, 0)), a, x, y, u.stream.get()));
        u.stream.sync();
        blas->flops += 2 * NLulong(std::max(n, 0));
    });
}

void d_dmul(NLBlas_t blas, int n, const double* a, const double* x, double* y) {
    guarded("nlHIPBlas()->Dmul", [&] {
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        NLHIP_CHECK(blas_dmul(size_t(std::max(n, 0)), a, x, y, u.stream.get()));
        u.stream.sync();
        blas->flops += NLulong(std::max(n, 0));
    });
}

void d_dzero(NLBlas_t, int n, double* x) {
    guarded("nlHIPBlas()->Dzero", [&] {
        Util& u = util();
        std::lock_guard<std::mutex> lock(u.mu);
        NLHIP_CHECK(fill<double>(size_t(std::max(n, 0)), 0.0, x, u.stream.get()));
        u.stream.sync();
    });
}

BlasImpl& host_blas() {
    static BlasImpl b = [] {
        BlasImpl r{};
        r.table.Malloc = b_malloc;
        r.table.Free = b_free;
        r.table.Memcpy = b_memcpy;
        r.table.Dcopy = h_dcopy;
        r.table.Dscal = h_dscal;
        r.table.Ddot = h_ddot;
        r.table.Dnrm2 = h_dnrm2;
        r.table.Daxpy = h_daxpy;
        r.table.Dmul = h_dmul;
        r.table.Dzero = h_dzero;
        r.table.has_unique_memory = NL_TRUE;
        r.device_vectors = false;
        return r;
    }();
    return b;
}

BlasImpl& hip_blas() {
    static BlasImpl b = [] {
        BlasImpl r{};
        r.table.Malloc = b_malloc;
        r.table.Free = b_free;
        r.table.Memcpy = b_memcpy;
        r.table.Dcopy = d_dcopy;
        r.table.Dscal = d_dscal;
        r.table.Ddot = h_ddot;   // staging is skipped for device pointers
        r.table.Dnrm2 = h_dnrm2;
        r.table.Daxpy = d_daxpy;
        r.table.Dmul = d_dmul;
        r.table.Dzero = d_dzero;
        r.table.has_unique_memory = NL_FALSE;
        r.device_vectors = true;
        return r;
    }();
    return b;
}

// --------------------------------------------------------------- context

enum class CtxState { Initial, System, Matrix, Row, MatrixDone, SystemDone, Solved };

struct Context {
    NLuint nb_variables = 0;
    NLuint nb_systems = 1;
    bool least_squares = false;
    bool symmetric = false;
    bool verbose = false;
    bool normalize_rows = false;
    NLenum solver = NL_SOLVER_DEFAULT;
    NLenum preconditioner = NL_PRECOND_JACOBI;
    double threshold = 1e-6;
    NLuint max_iterations = 0;  // 0: 5 x number of free variables
    NLuint inner_iterations = 5;
    double omega = 1.5;
    std::vector<double> values;   // nb_systems blocks of nb_variables
    std::vector<uint8_t> locked;
    CtxState state = CtxState::Initial;
    std::vector<NLuint> free_index;
    NLuint nb_free = 0;
    NLSparseMatrix* M = nullptr;
    std::vector<double> rhs;      // nb_systems blocks of nb_free
    NLuint current_row = 0;
    std::vector<std::pair<NLuint, double>> row;
    std::vector<double> row_rhs;
    double row_scaling = 1.0;
    // Least-squares rows recorded for the device assembly (scaled as
    // end_row scales them): entries over ls_rp, right-hand sides row-major
    // (nb_systems per row).  ls_device is false once the matrix went the
    // host way (an nlAddIJCoefficient in between); ls_sys is the assembled
    // system, ls_nnz its nonzeros.
    bool ls_device = false;
    std::vector<uint32_t> ls_rp, ls_var;
    std::vector<double> ls_coef, ls_rhs;
    std::unique_ptr<System> ls_sys;
    NLuint used_iterations = 0;
    double error = 0.0;
    double elapsed = 0.0;
    double gflops = 0.0;
    ~Context() {
        if (M) {
            M->destroy_func(reinterpret_cast<NLMatrix>(M));
            std::free(M);
        }
    }
};

Context* g_current = nullptr;

Context& cur(const char* fn) {
    if (!g_current) throw Error(NLHIP_ERR_INVALID, std::string(fn) + ": no current context");
    return *g_current;
}

void ensure_variables(Context& C) {
    const size_t sz = size_t(C.nb_variables) * C.nb_systems;
    if (C.values.size() != sz) C.values.assign(sz, 0.0);
    if (C.locked.size() != C.nb_variables) C.locked.assign(C.nb_variables, 0);
}

void check_var(const Context& C, NLuint i, const char* fn) {
    if (i >= C.nb_variables)
        throw Error(NLHIP_ERR_INVALID, std::string(fn) + ": variable index out of range");
}

void begin_matrix(Context& C) {
    ensure_variables(C);
    C.free_index.assign(C.nb_variables, ~NLuint(0));
    NLuint f = 0;
    for (NLuint i = 0; i < C.nb_variables; ++i)
        if (!C.locked[i]) C.free_index[i] = f++;
    C.nb_free = f;
    if (C.M) {
        C.M->destroy_func(reinterpret_cast<NLMatrix>(C.M));
        std::free(C.M);
        C.M = nullptr;
    }
    C.M = dyn_new(f, f);
    C.rhs.assign(size_t(f) * C.nb_systems, 0.0);
    C.current_row = 0;
    C.ls_device = C.least_squares;
    C.ls_rp.assign(1, 0);
    C.ls_var.clear();
    C.ls_coef.clear();
    C.ls_rhs.clear();
    C.ls_sys.reset();
}

// One row sum_k a_k x_{v_k} = rhs (coefficients and right-hand sides already
// scaled): locked variables move to the right-hand side in coefficient
// order, then the least-squares normal equations receive a_f1 a_f2 and
// a_f (rhs - locked part), or the row becomes row cur_row.
void host_row(Context& C, const std::vector<std::pair<NLuint, double>>& row, const double* rr) {
    std::vector<double> S(C.nb_systems);
    for (NLuint k = 0; k < C.nb_systems; ++k) {
        double s = rr[k];
        for (auto& c : row) {
            if (C.locked[c.first]) {
                double p = c.second * C.values[size_t(k) * C.nb_variables + c.first];
                s = s - p;
            }
        }
        S[k] = s;
    }
    if (C.least_squares) {
        for (auto& c1 : row) {
            const NLuint f1 = C.free_index[c1.first];
            if (f1 == ~NLuint(0)) continue;
            for (auto& c2 : row) {
                const NLuint f2 = C.free_index[c2.first];
                if (f2 == ~NLuint(0)) continue;
                double p = c1.second * c2.second;
                row_add(&C.M->row[f1], f2, p);
            }
            for (NLuint k = 0; k < C.nb_systems; ++k) {
                double q = c1.second * S[k];
                double& r = C.rhs[size_t(k) * C.nb_free + f1];
                r = r + q;
            }
        }
        C.M->version++;
    } else {
        const NLuint r = C.current_row;
        if (r >= C.nb_free) throw Error(NLHIP_ERR_INVALID, "nlEnd(NL_ROW): more rows than free variables");
        for (auto& c : row) {
            const NLuint f = C.free_index[c.first];
            if (f == ~NLuint(0)) continue;
            row_add(&C.M->row[r], f, c.second);
        }
        for (NLuint k = 0; k < C.nb_systems; ++k) C.rhs[size_t(k) * C.nb_free + r] = S[k];
        C.M->version++;
    }
}

// Row scaling (nlRowScaling, NL_NORMALIZE_ROWS), then the row is recorded
// for the device assembly (least squares) or added on the host.
void end_row(Context& C) {
    double w = C.row_scaling;
    if (C.normalize_rows) {
        double s2 = 0.0;
        for (auto& c : C.row) s2 += c.second * c.second;
        if (s2 > 0.0) w = w / std::sqrt(s2);
    }
    std::vector<std::pair<NLuint, double>> row = C.row;
    if (w != 1.0)
        for (auto& c : row) c.second = c.second * w;
    std::vector<double> rr(C.nb_systems);
    for (NLuint k = 0; k < C.nb_systems; ++k)
        rr[k] = (w != 1.0) ? C.row_rhs[k] * w : C.row_rhs[k];
    if (C.least_squares && C.ls_device) {
        if (C.ls_var.size() + row.size() > 0xffffffffull)
            throw Error(NLHIP_ERR_UNSUPPORTED, "more than 2^32 - 1 row coefficients");
        for (auto& c : row) {
            C.ls_var.push_back(c.first);
            C.ls_coef.push_back(c.second);
        }
        C.ls_rp.push_back(uint32_t(C.ls_var.size()));
        C.ls_rhs.insert(C.ls_rhs.end(), rr.begin(), rr.end());
        C.ls_sys.reset();
    } else {
        host_row(C, row, rr.data());
    }
    ++C.current_row;
}

// The recorded rows go through the host accumulation instead (a matrix that
// also receives nlAddIJCoefficient keeps OpenNL's order of additions).
void ls_to_host(Context& C) {
    if (!C.ls_device) return;
    C.ls_device = false;
    const size_t nrows = C.ls_rp.size() - 1;
    std::vector<std::pair<NLuint, double>> row;
    for (size_t r = 0; r < nrows; ++r) {
        row.clear();
        for (uint32_t e = C.ls_rp[r]; e < C.ls_rp[r + 1]; ++e)
            row.emplace_back(C.ls_var[e], C.ls_coef[e]);
        host_row(C, row, &C.ls_rhs[r * C.nb_systems]);
    }
    C.ls_rp.assign(1, 0);
    C.ls_var.clear();
    C.ls_coef.clear();
    C.ls_rhs.clear();
}

// The normal equations of the recorded rows, assembled on the device; the
// right-hand sides land in C.rhs.
System& ls_system(Context& C) {
    if (C.ls_sys) return *C.ls_sys;
    const uint32_t nrows = uint32_t(C.ls_rp.size() - 1);
    std::vector<double> rhs_t(size_t(C.nb_systems) * nrows);  // systems x rows
    for (uint32_t r = 0; r < nrows; ++r)
        for (NLuint k = 0; k < C.nb_systems; ++k)
            rhs_t[size_t(k) * nrows + r] = C.ls_rhs[size_t(r) * C.nb_systems + k];
    uint32_t nfree = 0;
    C.ls_sys.reset(ls_create(C.nb_variables, nrows, C.ls_rp.data(), C.ls_var.data(),
                             C.ls_coef.data(), C.nb_systems, rhs_t.data(), C.locked.data(),
                             C.values.data(), 0, C.rhs.data(), &nfree));
    return *C.ls_sys;
}

bool solve_context(Context& C) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!C.M) throw Error(NLHIP_ERR_INVALID, "nlSolve: no matrix (nlBegin(NL_MATRIX) missing)");
    const NLuint n = C.nb_free;
    C.used_iterations = 0;
    if (n == 0) {
        C.error = 0.0;
        return true;
    }
    std::unique_ptr<System> owned;
    System* S = nullptr;
    if (C.ls_device) {
        S = &ls_system(C);
    } else {
        std::vector<NLuint> rp, ci;
        std::vector<double> va;
        dyn_to_arrays(C.M, rp, ci, va);
        owned.reset(system_create(n, rp.data(), ci.data(), va.data(), 0));
        S = owned.get();
    }
    NLenum solver = C.solver;
    if (solver == NL_SOLVER_DEFAULT) solver = (C.least_squares || C.symmetric) ? NL_CG : NL_BICGSTAB;
    const NLuint max_iter = C.max_iterations ? C.max_iterations : 5 * n;
    std::vector<double> b(n), x(n);
    bool ok = true;
    // NL_PRECOND_SSOR: SSOR(NL_OMEGA) of the assembled matrix
    std::unique_ptr<Ssor, void (*)(Ssor*)> ssor(nullptr, ssor_destroy);
    if (C.preconditioner == NL_PRECOND_SSOR) ssor.reset(ssor_create(*S, C.omega));
    if (solver == NL_CG && C.nb_systems > 1 && C.nb_systems <= 32 && !ssor) {
        // NL_NB_SYSTEMS right-hand sides share every pass over the matrix
        const NLuint k = C.nb_systems;
        std::vector<double> X(size_t(n) * k);
        for (NLuint j = 0; j < k; ++j) {
            const double* vals = &C.values[size_t(j) * C.nb_variables];
            for (NLuint v = 0; v < C.nb_variables; ++v)
                if (C.free_index[v] != ~NLuint(0)) X[size_t(j) * n + C.free_index[v]] = vals[v];
        }
        nlhip_solve_params p;
        nlhipSolveParamsDefault(&p);
        p.eps = C.threshold;
        p.max_iter = max_iter;
        p.precond = (C.preconditioner == NL_PRECOND_JACOBI) ? NL_PRECOND_JACOBI : NL_PRECOND_NONE;
        std::vector<nlhip_solve_stats> per(k);
        nlhip_solve_stats tot;
        solve_multi(*S, k, C.rhs.data(), X.data(), p, per.data(), &tot);
        for (NLuint j = 0; j < k; ++j) {
            C.used_iterations += per[j].iterations;
            double* out = &C.values[size_t(j) * C.nb_variables];
            for (NLuint v = 0; v < C.nb_variables; ++v)
                if (C.free_index[v] != ~NLuint(0)) out[v] = X[size_t(j) * n + C.free_index[v]];
        }
        C.error = tot.rel_residual;
        C.elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return true;
    }
    for (NLuint k = 0; k < C.nb_systems; ++k) {
        const double* vals = &C.values[size_t(k) * C.nb_variables];
        for (NLuint v = 0; v < C.nb_variables; ++v)
            if (C.free_index[v] != ~NLuint(0)) x[C.free_index[v]] = vals[v];
        std::copy(C.rhs.begin() + size_t(k) * n, C.rhs.begin() + size_t(k + 1) * n, b.begin());
        if (solver == NL_CG && ssor) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b.data(), size_t(n) * 8, s);
            copy_in(S->ws64.x.get(), x.data(), size_t(n) * 8, s);
            double rel2 = 0.0;
            Precond pc;
            pc.ssor = ssor.get();
            C.used_iterations += pcg_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(), C.threshold,
                                            max_iter, &rel2);
            copy_out(x.data(), S->ws64.x.get(), size_t(n) * 8, s);
            S->stream.sync();
            C.error = std::sqrt(rel2);
        } else if (solver == NL_CG) {
            nlhip_solve_params p;
            nlhipSolveParamsDefault(&p);
            p.eps = C.threshold;
            p.max_iter = max_iter;
            p.precond = (C.preconditioner == NL_PRECOND_JACOBI) ? NL_PRECOND_JACOBI : NL_PRECOND_NONE;
            nlhip_solve_stats st;
            solve(*S, b.data(), x.data(), p, st);
            C.used_iterations += st.iterations;
            C.error = st.rel_residual;
        } else if (solver == NL_BICGSTAB) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b.data(), size_t(n) * 8, s);
            copy_in(S->ws64.x.get(), x.data(), size_t(n) * 8, s);
            double rel2 = 0.0;
            Precond pc;
            pc.dinv = (C.preconditioner == NL_PRECOND_JACOBI) ? S->dinv.get() : nullptr;
            pc.ssor = ssor.get();
            C.used_iterations += bicgstab_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(),
                                                 C.threshold, max_iter, &rel2);
            copy_out(x.data(), S->ws64.x.get(), size_t(n) * 8, s);
            S->stream.sync();
            C.error = std::sqrt(rel2);
        } else if (solver == NL_GMRES) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b.data(), size_t(n) * 8, s);
            copy_in(S->ws64.x.get(), x.data(), size_t(n) * 8, s);
            double rel2 = 0.0;
            Precond pc;
            pc.dinv = (C.preconditioner == NL_PRECOND_JACOBI) ? S->dinv.get() : nullptr;
            pc.ssor = ssor.get();
            C.used_iterations += gmres_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(),
                                              C.threshold, max_iter,
                                              C.inner_iterations ? C.inner_iterations : 5, &rel2);
            copy_out(x.data(), S->ws64.x.get(), size_t(n) * 8, s);
            S->stream.sync();
            C.error = std::sqrt(rel2);
        } else {
            throw Error(NLHIP_ERR_UNSUPPORTED, "nlSolve: solver not available on the GPU");
        }
        double* out = &C.values[size_t(k) * C.nb_variables];
        for (NLuint v = 0; v < C.nb_variables; ++v)
            if (C.free_index[v] != ~NLuint(0)) out[v] = x[C.free_index[v]];
    }
    C.elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (C.verbose) {
        std::fprintf(stderr, "[NL] solved %u unknowns, %u iterations, error %.3e, %.3f s (GPU)\n",
                     n, C.used_iterations, C.error, C.elapsed);
    }
    return ok;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

// ---- matrices ---------------------------------------------------------

NLAPI NLMatrix NLAPIENTRY nlSparseMatrixNew(NLuint m, NLuint n, NLenum storage) {
    NLMatrix r = nullptr;
    guarded("nlSparseMatrixNew", [&] {
        if (storage != NL_MATRIX_STORE_ROWS)
            throw Error(NLHIP_ERR_UNSUPPORTED, "only NL_MATRIX_STORE_ROWS storage is supported");
        NLSparseMatrix* M = dyn_new(m, n);
        M->storage = storage;
        r = reinterpret_cast<NLMatrix>(M);
    });
    return r;
}

NLAPI void NLAPIENTRY nlSparseMatrixAdd(NLSparseMatrix* M, NLuint i, NLuint j, NLdouble value) {
    guarded("nlSparseMatrixAdd", [&] {
        if (!M || M->type != NL_MATRIX_SPARSE_DYNAMIC)
            throw Error(NLHIP_ERR_INVALID, "not a dynamic matrix");
        if (i >= M->m || j >= M->n) throw Error(NLHIP_ERR_INVALID, "index out of bounds");
        row_add(&M->row[i], j, value);
        M->version++;
    });
}

NLAPI void NLAPIENTRY nlSparseMatrixZero(NLSparseMatrix* M) {
    if (!M || M->type != NL_MATRIX_SPARSE_DYNAMIC) return;
    for (NLuint i = 0; i < M->m; ++i) M->row[i].size = 0;
    M->version++;
}

NLAPI NLuint NLAPIENTRY nlSparseMatrixNNZ(NLSparseMatrix* M) {
    if (!M || M->type != NL_MATRIX_SPARSE_DYNAMIC) return 0;
    uint64_t s = 0;
    for (NLuint i = 0; i < M->m; ++i) s += M->row[i].size;
    return NLuint(s);
}

NLAPI NLMatrix NLAPIENTRY nlCRSMatrixNewFromArrays(NLuint m, NLuint n, const NLuint* rowptr,
                                                   const NLuint* colind, const NLdouble* val) {
    NLMatrix r = nullptr;
    guarded("nlCRSMatrixNewFromArrays", [&] {
        if (!rowptr) throw Error(NLHIP_ERR_INVALID, "null rowptr");
        r = reinterpret_cast<NLMatrix>(crs_from_arrays(m, n, rowptr, colind, val));
    });
    return r;
}

NLAPI void NLAPIENTRY nlMatrixCompress(NLMatrix* M) {
    guarded("nlMatrixCompress", [&] {
        if (!M || !*M) throw Error(NLHIP_ERR_INVALID, "null matrix");
        if ((*M)->type != NL_MATRIX_SPARSE_DYNAMIC) return;
        NLSparseMatrix* D = reinterpret_cast<NLSparseMatrix*>(*M);
        std::vector<NLuint> rp, ci;
        std::vector<double> va;
        dyn_to_arrays(D, rp, ci, va);
        NLCRSMatrix* C = crs_from_arrays(D->m, D->n, rp.data(), ci.data(), va.data());
        nlDeleteMatrix(*M);
        *M = reinterpret_cast<NLMatrix>(C);
    });
}

NLAPI void NLAPIENTRY nlMultMatrixVector(NLMatrix M, const double* x, double* y) {
    if (!M) {
        log_error("nlMultMatrixVector: null matrix");
        return;
    }
    M->mult_func(M, x, y);
}

NLAPI void NLAPIENTRY nlDeleteMatrix(NLMatrix M) {
    if (!M) return;
    M->destroy_func(M);
    if (M->type == NL_MATRIX_OTHER) {
        HIPMatrix* H = reinterpret_cast<HIPMatrix*>(M);
        if (H->magic == kMagicHIP) {
            delete H;
            return;
        }
        JacobiMatrix* J = reinterpret_cast<JacobiMatrix*>(M);
        if (J->magic == kMagicJacobi) {
            delete J;
            return;
        }
        if (SSORMatrix* Q = as_ssor(M)) {
            delete Q;
            return;
        }
    }
    std::free(M);
}

NLAPI NLMatrix NLAPIENTRY nlMatrixFactorize(NLMatrix, NLenum) {
    log_error("nlMatrixFactorize: direct solvers (SUPERLU, CHOLMOD) are not part of this library");
    return nullptr;
}

NLAPI NLMatrix NLAPIENTRY nlNewJacobiPreconditioner(NLMatrix M) {
    NLMatrix r = nullptr;
    guarded("nlNewJacobiPreconditioner", [&] { r = jacobi_new(M); });
    return r;
}

NLAPI NLMatrix NLAPIENTRY nlNewSSORPreconditioner(NLMatrix M, NLdouble omega) {
    NLMatrix r = nullptr;
    guarded("nlNewSSORPreconditioner", [&] { r = ssor_new(M, omega); });
    return r;
}

// ---- BLAS ---------------------------------------------------------------

NLAPI NLBlas_t NLAPIENTRY nlHostBlas(void) { return &host_blas().table; }
NLAPI NLBlas_t NLAPIENTRY nlHIPBlas(void) { return &hip_blas().table; }

NLAPI void NLAPIENTRY nlBlasResetStats(NLBlas_t blas) {
    if (!blas) return;
    blas->flops = 0;
    blas->start_time = NLulong(std::chrono::duration_cast<std::chrono::microseconds>(
                                   std::chrono::steady_clock::now().time_since_epoch())
                                   .count());
}

NLAPI double NLAPIENTRY nlBlasGFlops(NLBlas_t blas) {
    if (!blas) return 0.0;
    const NLulong now = NLulong(std::chrono::duration_cast<std::chrono::microseconds>(
                                    std::chrono::steady_clock::now().time_since_epoch())
                                    .count());
    const double s = 1e-6 * double(now - blas->start_time);
    return s > 0 ? 1e-9 * double(blas->flops) / s : 0.0;
}

// ---- iterative solver ---------------------------------------------------

NLAPI NLuint NLAPIENTRY nlSolveSystemIterative(NLBlas_t blas, NLMatrix M, NLMatrix P,
                                               NLdouble* b, NLdouble* x, NLenum solver,
                                               double eps, NLuint max_iter, NLuint inner_iter) {
    NLuint its = 0;
    guarded("nlSolveSystemIterative", [&] {
        if (!b || !x) throw Error(NLHIP_ERR_INVALID, "null b or x");
        System* S = system_of(M);
        const Precond pc = precond_of(P, S->n);
        const double* dinv = pc.dinv;
        if ((solver == NL_CG || solver == NL_SOLVER_DEFAULT) && pc.ssor) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(S->n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b, size_t(S->n) * 8, s);
            copy_in(S->ws64.x.get(), x, size_t(S->n) * 8, s);
            its = pcg_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(), eps, max_iter, nullptr);
            copy_out(x, S->ws64.x.get(), size_t(S->n) * 8, s);
            S->stream.sync();
            if (blas) blas->flops += NLulong(its) * (4 * S->nnz + 12 * NLulong(S->n));
        } else if (solver == NL_CG || solver == NL_SOLVER_DEFAULT) {
            nlhip_solve_params p;
            nlhipSolveParamsDefault(&p);
            p.eps = eps;
            p.max_iter = max_iter;
            p.precond = dinv ? NL_PRECOND_JACOBI : NL_PRECOND_NONE;
            nlhip_solve_stats st;
            solve(*S, b, x, p, st, dinv);
            its = st.iterations;
            if (blas) blas->flops += NLulong(its) * (2 * S->nnz + 12 * NLulong(S->n));
        } else if (solver == NL_BICGSTAB) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(S->n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b, size_t(S->n) * 8, s);
            copy_in(S->ws64.x.get(), x, size_t(S->n) * 8, s);
            its = bicgstab_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(), eps, max_iter,
                                  nullptr);
            copy_out(x, S->ws64.x.get(), size_t(S->n) * 8, s);
            S->stream.sync();
            if (blas) blas->flops += NLulong(its) * (4 * S->nnz + 20 * NLulong(S->n));
        } else if (solver == NL_GMRES) {
            std::lock_guard<std::mutex> lock(S->mu);
            S->ws64.ensure(S->n, S->grid);
            hipStream_t s = S->stream.get();
            copy_in(S->ws64.b.get(), b, size_t(S->n) * 8, s);
            copy_in(S->ws64.x.get(), x, size_t(S->n) * 8, s);
            const NLuint m = inner_iter ? inner_iter : 5;  // matrix.cpp:385 passes 5
            its = gmres_device(*S, pc, S->ws64.b.get(), S->ws64.x.get(), eps, max_iter, m,
                               nullptr);
            copy_out(x, S->ws64.x.get(), size_t(S->n) * 8, s);
            S->stream.sync();
            if (blas) blas->flops += NLulong(its) * (2 * S->nnz + NLulong(2 * m + 8) * S->n);
        } else {
            throw Error(NLHIP_ERR_UNSUPPORTED,
                        "solver not available on the GPU (CG, BiCGSTAB, GMRES)");
        }
    });
    return its;
}

// ---- extensions ---------------------------------------------------------

NLAPI NLboolean NLAPIENTRY nlInitExtension_HIP(void) {
    try {
        device();
        return NL_TRUE;
    } catch (const Error& e) {
        log_error(std::string("nlInitExtension(\"HIP\"): ") + e.what());
        return NL_FALSE;
    }
}

NLAPI NLboolean NLAPIENTRY nlExtensionIsInitialized_HIP(void) {
    return device_usable() ? NL_TRUE : NL_FALSE;
}

NLAPI NLboolean NLAPIENTRY nlInitExtension(const char* extension) {
    if (!extension) return NL_FALSE;
    if (std::strcmp(extension, "HIP") == 0) return nlInitExtension_HIP();
    return NL_FALSE;
}

NLAPI NLboolean NLAPIENTRY nlExtensionIsInitialized(const char* extension) {
    if (!extension) return NL_FALSE;
    if (std::strcmp(extension, "HIP") == 0) return nlExtensionIsInitialized_HIP();
    return NL_FALSE;
}

NLAPI NLMatrix NLAPIENTRY nlHIPMatrixNewFromCRSMatrix(NLMatrix M) {
    NLMatrix r = nullptr;
    guarded("nlHIPMatrixNewFromCRSMatrix", [&] {
        if (!M) throw Error(NLHIP_ERR_INVALID, "null matrix");
        std::unique_ptr<System> S;
        if (M->type == NL_MATRIX_CRS) {
            NLCRSMatrix* C = reinterpret_cast<NLCRSMatrix*>(M);
            if (C->m != C->n) throw Error(NLHIP_ERR_UNSUPPORTED, "non-square matrix");
            S.reset(system_create(C->m, C->rowptr, C->colind, C->val, 0));
        } else if (M->type == NL_MATRIX_SPARSE_DYNAMIC) {
            NLSparseMatrix* D = reinterpret_cast<NLSparseMatrix*>(M);
            if (D->m != D->n) throw Error(NLHIP_ERR_UNSUPPORTED, "non-square matrix");
            std::vector<NLuint> rp, ci;
            std::vector<double> va;
            dyn_to_arrays(D, rp, ci, va);
            S.reset(system_create(D->m, rp.data(), ci.data(), va.data(), 0));
        } else {
            throw Error(NLHIP_ERR_UNSUPPORTED, "expected a CRS or dynamic matrix");
        }
        auto H = std::make_unique<HIPMatrix>();
        H->m = H->n = S->n;
        H->type = NL_MATRIX_OTHER;
        H->destroy_func = hip_destroy;
        H->mult_func = hip_mult;
        H->magic = kMagicHIP;
        H->sys = S.release();
        r = reinterpret_cast<NLMatrix>(H.release());
    });
    return r;
}

NLAPI NLMatrix NLAPIENTRY nlHIPJacobiPreconditionerNewFromCRSMatrix(NLMatrix M) {
    NLMatrix r = nullptr;
    guarded("nlHIPJacobiPreconditionerNewFromCRSMatrix", [&] { r = jacobi_new(M); });
    return r;
}

// ---- context API --------------------------------------------------------

NLAPI NLContext NLAPIENTRY nlNewContext(void) {
    Context* C = new Context();
    g_current = C;
    return C;
}

NLAPI void NLAPIENTRY nlDeleteContext(NLContext context) {
    Context* C = static_cast<Context*>(context);
    if (g_current == C) g_current = nullptr;
    delete C;
}

NLAPI void NLAPIENTRY nlMakeCurrent(NLContext context) { g_current = static_cast<Context*>(context); }
NLAPI NLContext NLAPIENTRY nlGetCurrent(void) { return g_current; }

NLAPI void NLAPIENTRY nlSolverParameterd(NLenum pname, NLdouble param) {
    guarded("nlSolverParameterd", [&] {
        Context& C = cur("nlSolverParameterd");
        switch (pname) {
            case NL_THRESHOLD: C.threshold = param; break;
            case NL_OMEGA: C.omega = param; break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown parameter");
        }
    });
}

NLAPI void NLAPIENTRY nlSolverParameteri(NLenum pname, NLint param) {
    guarded("nlSolverParameteri", [&] {
        Context& C = cur("nlSolverParameteri");
        if (C.state != CtxState::Initial && pname == NL_NB_VARIABLES)
            throw Error(NLHIP_ERR_INVALID, "NL_NB_VARIABLES set after nlBegin(NL_SYSTEM)");
        switch (pname) {
            case NL_SOLVER: C.solver = NLenum(param); break;
            case NL_NB_VARIABLES:
                if (param < 0) throw Error(NLHIP_ERR_INVALID, "negative NL_NB_VARIABLES");
                C.nb_variables = NLuint(param);
                break;
            case NL_NB_SYSTEMS:
                if (param < 1) throw Error(NLHIP_ERR_INVALID, "NL_NB_SYSTEMS must be >= 1");
                C.nb_systems = NLuint(param);
                break;
            case NL_LEAST_SQUARES: C.least_squares = param != 0; break;
            case NL_MAX_ITERATIONS: C.max_iterations = NLuint(std::max(param, 0)); break;
            case NL_SYMMETRIC: C.symmetric = param != 0; break;
            case NL_INNER_ITERATIONS: C.inner_iterations = NLuint(std::max(param, 0)); break;
            case NL_PRECONDITIONER: C.preconditioner = NLenum(param); break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown parameter");
        }
    });
}

NLAPI void NLAPIENTRY nlGetBooleanv(NLenum pname, NLboolean* params) {
    guarded("nlGetBooleanv", [&] {
        Context& C = cur("nlGetBooleanv");
        switch (pname) {
            case NL_LEAST_SQUARES: *params = C.least_squares; break;
            case NL_SYMMETRIC: *params = C.symmetric; break;
            case NL_VERBOSE: *params = C.verbose; break;
            case NL_NORMALIZE_ROWS: *params = C.normalize_rows; break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown parameter");
        }
    });
}

NLAPI void NLAPIENTRY nlGetDoublev(NLenum pname, NLdouble* params) {
    guarded("nlGetDoublev", [&] {
        Context& C = cur("nlGetDoublev");
        switch (pname) {
            case NL_THRESHOLD: *params = C.threshold; break;
            case NL_OMEGA: *params = C.omega; break;
            case NL_ERROR: *params = C.error; break;
            case NL_ELAPSED_TIME: *params = C.elapsed; break;
            case NL_GFLOPS: *params = C.gflops; break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown parameter");
        }
    });
}

NLAPI void NLAPIENTRY nlGetIntegerv(NLenum pname, NLint* params) {
    guarded("nlGetIntegerv", [&] {
        Context& C = cur("nlGetIntegerv");
        switch (pname) {
            case NL_SOLVER: *params = NLint(C.solver); break;
            case NL_NB_VARIABLES: *params = NLint(C.nb_variables); break;
            case NL_NB_SYSTEMS: *params = NLint(C.nb_systems); break;
            case NL_LEAST_SQUARES: *params = C.least_squares; break;
            case NL_MAX_ITERATIONS: *params = NLint(C.max_iterations); break;
            case NL_SYMMETRIC: *params = C.symmetric; break;
            case NL_USED_ITERATIONS: *params = NLint(C.used_iterations); break;
            case NL_PRECONDITIONER: *params = NLint(C.preconditioner); break;
            case NL_INNER_ITERATIONS: *params = NLint(C.inner_iterations); break;
            case NL_NNZ:
                if (C.ls_device && C.nb_free) *params = NLint(ls_system(C).nnz);
                else *params = C.M ? NLint(nlSparseMatrixNNZ(C.M)) : 0;
                break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown parameter");
        }
    });
}

NLAPI void NLAPIENTRY nlEnable(NLenum pname) {
    guarded("nlEnable", [&] {
        Context& C = cur("nlEnable");
        if (pname == NL_VERBOSE) C.verbose = true;
        else if (pname == NL_NORMALIZE_ROWS) C.normalize_rows = true;
        else if (pname == NL_NO_VARIABLES_INDIRECTION) {
        } else throw Error(NLHIP_ERR_INVALID, "unknown parameter");
    });
}

NLAPI void NLAPIENTRY nlDisable(NLenum pname) {
    guarded("nlDisable", [&] {
        Context& C = cur("nlDisable");
        if (pname == NL_VERBOSE) C.verbose = false;
        else if (pname == NL_NORMALIZE_ROWS) C.normalize_rows = false;
        else if (pname == NL_NO_VARIABLES_INDIRECTION) {
        } else throw Error(NLHIP_ERR_INVALID, "unknown parameter");
    });
}

NLAPI NLboolean NLAPIENTRY nlIsEnabled(NLenum pname) {
    NLboolean r = NL_FALSE;
    guarded("nlIsEnabled", [&] {
        Context& C = cur("nlIsEnabled");
        if (pname == NL_VERBOSE) r = C.verbose;
        else if (pname == NL_NORMALIZE_ROWS) r = C.normalize_rows;
    });
    return r;
}

NLAPI void NLAPIENTRY nlSetVariable(NLuint i, NLdouble value) { nlMultiSetVariable(i, 0, value); }

NLAPI void NLAPIENTRY nlMultiSetVariable(NLuint i, NLuint k, NLdouble value) {
    guarded("nlSetVariable", [&] {
        Context& C = cur("nlSetVariable");
        ensure_variables(C);
        check_var(C, i, "nlSetVariable");
        if (k >= C.nb_systems) throw Error(NLHIP_ERR_INVALID, "system index out of range");
        C.values[size_t(k) * C.nb_variables + i] = value;
    });
}

NLAPI NLdouble NLAPIENTRY nlGetVariable(NLuint i) { return nlMultiGetVariable(i, 0); }

NLAPI NLdouble NLAPIENTRY nlMultiGetVariable(NLuint i, NLuint k) {
    double r = 0.0;
    guarded("nlGetVariable", [&] {
        Context& C = cur("nlGetVariable");
        ensure_variables(C);
        check_var(C, i, "nlGetVariable");
        if (k >= C.nb_systems) throw Error(NLHIP_ERR_INVALID, "system index out of range");
        r = C.values[size_t(k) * C.nb_variables + i];
    });
    return r;
}

NLAPI void NLAPIENTRY nlLockVariable(NLuint index) {
    guarded("nlLockVariable", [&] {
        Context& C = cur("nlLockVariable");
        ensure_variables(C);
        check_var(C, index, "nlLockVariable");
        if (C.state == CtxState::Matrix || C.state == CtxState::Row)
            throw Error(NLHIP_ERR_INVALID, "nlLockVariable inside nlBegin(NL_MATRIX)");
        C.locked[index] = 1;
    });
}

NLAPI void NLAPIENTRY nlUnlockVariable(NLuint index) {
    guarded("nlUnlockVariable", [&] {
        Context& C = cur("nlUnlockVariable");
        ensure_variables(C);
        check_var(C, index, "nlUnlockVariable");
        if (C.state == CtxState::Matrix || C.state == CtxState::Row)
            throw Error(NLHIP_ERR_INVALID, "nlUnlockVariable inside nlBegin(NL_MATRIX)");
        C.locked[index] = 0;
    });
}

NLAPI NLboolean NLAPIENTRY nlVariableIsLocked(NLuint index) {
    NLboolean r = NL_FALSE;
    guarded("nlVariableIsLocked", [&] {
        Context& C = cur("nlVariableIsLocked");
        ensure_variables(C);
        check_var(C, index, "nlVariableIsLocked");
        r = C.locked[index] ? NL_TRUE : NL_FALSE;
    });
    return r;
}

NLAPI void NLAPIENTRY nlBegin(NLenum prim) {
    guarded("nlBegin", [&] {
        Context& C = cur("nlBegin");
        switch (prim) {
            case NL_SYSTEM:
                ensure_variables(C);
                C.state = CtxState::System;
                break;
            case NL_MATRIX:
                if (C.state != CtxState::System && C.state != CtxState::MatrixDone &&
                    C.state != CtxState::Solved && C.state != CtxState::SystemDone)
                    throw Error(NLHIP_ERR_INVALID, "nlBegin(NL_MATRIX) outside NL_SYSTEM");
                begin_matrix(C);
                C.state = CtxState::Matrix;
                break;
            case NL_ROW:
                if (C.state != CtxState::Matrix)
                    throw Error(NLHIP_ERR_INVALID, "nlBegin(NL_ROW) outside NL_MATRIX");
                C.row.clear();
                C.row_rhs.assign(C.nb_systems, 0.0);
                C.row_scaling = 1.0;
                C.state = CtxState::Row;
                break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown primitive");
        }
    });
}

NLAPI void NLAPIENTRY nlEnd(NLenum prim) {
    guarded("nlEnd", [&] {
        Context& C = cur("nlEnd");
        switch (prim) {
            case NL_SYSTEM: C.state = CtxState::SystemDone; break;
            case NL_MATRIX:
                if (C.state != CtxState::Matrix)
                    throw Error(NLHIP_ERR_INVALID, "nlEnd(NL_MATRIX) without nlBegin(NL_MATRIX)");
                C.state = CtxState::MatrixDone;
                break;
            case NL_ROW:
                if (C.state != CtxState::Row)
                    throw Error(NLHIP_ERR_INVALID, "nlEnd(NL_ROW) without nlBegin(NL_ROW)");
                end_row(C);
                C.state = CtxState::Matrix;
                break;
            default: throw Error(NLHIP_ERR_INVALID, "unknown primitive");
        }
    });
}

NLAPI void NLAPIENTRY nlCoefficient(NLuint i, NLdouble value) {
    guarded("nlCoefficient", [&] {
        Context& C = cur("nlCoefficient");
        if (C.state != CtxState::Row) throw Error(NLHIP_ERR_INVALID, "nlCoefficient outside NL_ROW");
        check_var(C, i, "nlCoefficient");
        C.row.emplace_back(i, value);
    });
}

NLAPI void NLAPIENTRY nlAddIJCoefficient(NLuint i, NLuint j, NLdouble value) {
    guarded("nlAddIJCoefficient", [&] {
        Context& C = cur("nlAddIJCoefficient");
        if (C.state != CtxState::Matrix)
            throw Error(NLHIP_ERR_INVALID, "nlAddIJCoefficient outside NL_MATRIX");
        check_var(C, i, "nlAddIJCoefficient");
        check_var(C, j, "nlAddIJCoefficient");
        ls_to_host(C);
        const NLuint fi = C.free_index[i], fj = C.free_index[j];
        if (fi == ~NLuint(0)) return;
        if (fj == ~NLuint(0)) {
            for (NLuint k = 0; k < C.nb_systems; ++k) {
                double p = value * C.values[size_t(k) * C.nb_variables + j];
                double& r = C.rhs[size_t(k) * C.nb_free + fi];
                r = r - p;
            }
        } else {
            row_add(&C.M->row[fi], fj, value);
            C.M->version++;
        }
    });
}

NLAPI void NLAPIENTRY nlRightHandSide(NLdouble value) { nlMultiRightHandSide(0, value); }

NLAPI void NLAPIENTRY nlMultiRightHandSide(NLuint k, NLdouble value) {
    guarded("nlRightHandSide", [&] {
        Context& C = cur("nlRightHandSide");
        if (C.state != CtxState::Row) throw Error(NLHIP_ERR_INVALID, "nlRightHandSide outside NL_ROW");
        if (k >= C.nb_systems) throw Error(NLHIP_ERR_INVALID, "system index out of range");
        C.row_rhs[k] = value;
    });
}

NLAPI void NLAPIENTRY nlRowScaling(NLdouble value) {
    guarded("nlRowScaling", [&] {
        Context& C = cur("nlRowScaling");
        if (C.state != CtxState::Row) throw Error(NLHIP_ERR_INVALID, "nlRowScaling outside NL_ROW");
        C.row_scaling = value;
    });
}

NLAPI NLboolean NLAPIENTRY nlSolve(void) {
    NLboolean ok = NL_FALSE;
    guarded("nlSolve", [&] {
        Context& C = cur("nlSolve");
        if (C.state == CtxState::Row || C.state == CtxState::Matrix)
            throw Error(NLHIP_ERR_INVALID, "nlSolve inside nlBegin(NL_MATRIX)");
        ok = solve_context(C) ? NL_TRUE : NL_FALSE;
        C.state = CtxState::Solved;
    });
    return ok;
}

}  // extern "C"
