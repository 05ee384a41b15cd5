// Compile-level drop-in check: the calls Graphite's NL wrapper and the
// WarpDrive plugin make into OpenNL, in the same call shapes, compiled
// against include/nl.h and linked to libnlhip.so without any change to the
// calling code:
//   * NLBlas member calls through the table, impl_->Dcopy(impl_, n, x, 1,
//     y, 1) etc. (src/lib/OGF/scene_graph/NL/blas.cpp:65-148);
//   * the dynamic matrix: nlSparseMatrixNew(m, n, NL_MATRIX_STORE_ROWS),
//     nlSparseMatrixAdd((NLSparseMatrix*)(impl_), i, j, a), impl_->m / n /
//     type, nlMatrixCompress(&impl_), nlMultMatrixVector, nlMatrixFactorize,
//     nlDeleteMatrix (src/lib/OGF/scene_graph/NL/matrix.cpp:121-143,163,
//     296-336);
//   * solve_iterative's nlNewJacobiPreconditioner / nlSolveSystemIterative(
//     nlHostBlas(), impl_, P, b, x, NL_CG, eps, max_iter, 5) / nlDeleteMatrix
//     (matrix.cpp:340-398);
//   * the voxel interpolation context sequence (plugins/OGF/WarpDrive/
//     commands/voxel_grob_transport_commands.cpp:231-287).
// Usage: call_shapes host   (matrix building only: runs without a GPU)
//        call_shapes gpu    (everything; nlInitExtension("HIP") first)
// Prints one result per line, doubles as C99 hex floats.
#include <cstdio>
#include <cstring>
#include <vector>

#include "nl.h"

namespace {

// The wrapper's Matrix, reduced to the calls it makes.
struct Matrix {
    NLMatrix impl_;
    Matrix(NLuint m, NLuint n) { impl_ = nlSparseMatrixNew(m, n, NL_MATRIX_STORE_ROWS); }
    ~Matrix() {
        nlDeleteMatrix(impl_);
        impl_ = nullptr;
    }
    NLuint get_m() const { return NLuint(impl_->m); }
    NLuint get_n() const { return NLuint(impl_->n); }
    const char* format() const {
        if (impl_->type == NL_MATRIX_SPARSE_DYNAMIC) return "Dynamic";
        if (impl_->type == NL_MATRIX_CRS) return "CRS";
        return "Factorized";
    }
    bool add_coefficient(NLuint i, NLuint j, double a) {
        if (impl_->type != NL_MATRIX_SPARSE_DYNAMIC) return false;
        if (i >= NLuint(impl_->m) || j >= NLuint(impl_->n)) return false;
        nlSparseMatrixAdd((NLSparseMatrix*)(impl_), i, j, a);
        return true;
    }
    void mult(const double* x, double* y) const { nlMultMatrixVector(impl_, x, y); }
    void compress() { nlMatrixCompress(&impl_); }
    NLMatrix factorize() const { return nlMatrixFactorize(impl_, NL_SUPERLU_EXT); }
    NLuint solve_iterative(const double* b, double* x, NLenum nl_solver, bool jacobi,
                           NLuint max_iter, double eps) {
        NLMatrix P = nullptr;
        if (jacobi) P = nlNewJacobiPreconditioner(impl_);
        NLuint inner_iter = 5;
        std::memset(x, 0, sizeof(double) * get_n());
        const NLuint its = nlSolveSystemIterative(nlHostBlas(), impl_, P, const_cast<double*>(b),
                                                  x, nl_solver, eps, max_iter, inner_iter);
        if (P != nullptr) nlDeleteMatrix(P);
        return its;
    }
};

// The wrapper's Blas, reduced to the calls it makes.
struct Blas {
    NLBlas_t impl_;
    Blas() { impl_ = nlHostBlas(); }
    void copy(const std::vector<double>& x, std::vector<double>& y) {
        NLint n = NLint(x.size());
        impl_->Dcopy(impl_, n, x.data(), 1, y.data(), 1);
    }
    void scal(double a, std::vector<double>& x) {
        NLint n = NLint(x.size());
        impl_->Dscal(impl_, n, a, x.data(), 1);
    }
    double dot(const std::vector<double>& x, const std::vector<double>& y) {
        NLint n = NLint(x.size());
        return impl_->Ddot(impl_, n, x.data(), 1, y.data(), 1);
    }
    double nrm2(const std::vector<double>& x) {
        NLint n = NLint(x.size());
        return impl_->Dnrm2(impl_, n, x.data(), 1);
    }
    void axpy(double a, const std::vector<double>& x, std::vector<double>& y) {
        NLint n = NLint(x.size());
        impl_->Daxpy(impl_, n, a, x.data(), 1, y.data(), 1);
    }
};

void print_vec(const char* key, const double* v, size_t n) {
    std::printf("%s", key);
    for (size_t i = 0; i < n; ++i) std::printf(" %a", v[i]);
    std::printf("\n");
}

// A 5 x 4 x 3 periodic voxel grid: locked cells keep their value, the others
// are interpolated by the least-squares 7-point Laplacian rows.
constexpr NLuint kNu = 5, kNv = 4, kNw = 3;
NLuint linear_index(NLuint u, NLuint v, NLuint w) { return (w * kNv + v) * kNu + u; }
double attribute_of(NLuint i) { return double((i * 37u) % 23u) / 8.0; }
bool interpolated(NLuint i) { return (i % 3u) != 0u; }

void voxel_context() {
    const NLuint nu = kNu, nv = kNv, nw = kNw;
    nlNewContext();
    nlSolverParameteri(NL_LEAST_SQUARES, NL_TRUE);
    nlSolverParameteri(NL_NB_VARIABLES, NLint(nu * nv * nw));
    nlSolverParameterd(NL_THRESHOLD, 1e-12);
    nlSolverParameteri(NL_MAX_ITERATIONS, 10000);
    nlBegin(NL_SYSTEM);
    for (NLuint i = 0; i < nu * nv * nw; ++i) {
        nlSetVariable(i, attribute_of(i));
        if (!interpolated(i)) nlLockVariable(i);
    }
    nlBegin(NL_MATRIX);
    for (NLuint w = 0; w < nw; ++w) {
        for (NLuint v = 0; v < nv; ++v) {
            for (NLuint u = 0; u < nu; ++u) {
                NLuint um = (u + nu - 1) % nu, up = (u + 1) % nu;
                NLuint vm = (v + nv - 1) % nv, vp = (v + 1) % nv;
                NLuint wm = (w + nw - 1) % nw, wp = (w + 1) % nw;
                nlBegin(NL_ROW);
                nlCoefficient(linear_index(u, v, w), -6.0);
                nlCoefficient(linear_index(um, v, w), 1.0);
                nlCoefficient(linear_index(up, v, w), 1.0);
                nlCoefficient(linear_index(u, vm, w), 1.0);
                nlCoefficient(linear_index(u, vp, w), 1.0);
                nlCoefficient(linear_index(u, v, wm), 1.0);
                nlCoefficient(linear_index(u, v, wp), 1.0);
                nlEnd(NL_ROW);
            }
        }
    }
    nlEnd(NL_MATRIX);
    nlEnd(NL_SYSTEM);
    const NLboolean ok = nlSolve();
    std::vector<double> out(nu * nv * nw);
    for (NLuint i = 0; i < nu * nv * nw; ++i) out[i] = nlGetVariable(i);
    NLint used = 0;
    nlGetIntegerv(NL_USED_ITERATIONS, &used);
    nlDeleteContext(nlGetCurrent());
    std::printf("context_ok %d\ncontext_iterations %d\n", int(ok), used);
    print_vec("context_x", out.data(), out.size());
}

}  // namespace

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    if (gpu && !nlInitExtension("HIP")) {
        std::printf("init_failed\n");
        return 2;
    }
    // the matrix_kat of the reference's linear_algebra.lua, symmetric case,
    // with one duplicate add to show the accumulation
    Matrix M(2, 2);
    std::printf("shape %u %u %s\n", M.get_m(), M.get_n(), M.format());
    M.add_coefficient(0, 0, 0.5);
    M.add_coefficient(0, 0, 0.5);
    M.add_coefficient(0, 1, 2.0);
    M.add_coefficient(1, 0, 2.0);
    M.add_coefficient(1, 1, 3.0);
    std::printf("add_out_of_bounds %d\n", int(M.add_coefficient(2, 0, 1.0)));
    std::printf("factorize_null %d\n", int(M.factorize() == nullptr));
    Matrix D(3, 3);  // compressed: rows reach the CRS in column order
    D.add_coefficient(2, 1, 4.0);
    D.add_coefficient(0, 2, -1.0);
    D.add_coefficient(0, 0, 2.0);
    D.add_coefficient(2, 1, 0.25);
    D.add_coefficient(1, 1, 5.0);
    D.compress();
    std::printf("compressed_format %s\n", D.format());
    const NLCRSMatrix* C = (const NLCRSMatrix*)(D.impl_);
    std::printf("crs_rowptr %u %u %u %u\n", C->rowptr[0], C->rowptr[1], C->rowptr[2],
                C->rowptr[3]);
    std::printf("crs_colind");
    for (NLuint j = 0; j < C->rowptr[3]; ++j) std::printf(" %u", C->colind[j]);
    std::printf("\n");
    print_vec("crs_val", C->val, C->rowptr[3]);
    if (!gpu) return 0;

    // products and solves (device)
    const double xin[3] = {1.0, 2.0, 3.0};
    double yout[3] = {0, 0, 0};
    D.mult(xin, yout);
    print_vec("mult", yout, 3);
    const double b[2] = {5.0, 6.0};
    double x[2] = {0, 0};
    const NLuint its = M.solve_iterative(b, x, NL_CG, true, 10, 1e-12);
    std::printf("cg_iterations %u\n", its);
    print_vec("cg_x", x, 2);
    Blas blas;
    std::vector<double> u = {1.0, 2.0, 3.0, 4.0, 0.5}, v = {0.25, -1.0, 2.0, 8.0, 16.0};
    std::vector<double> w(5, 0.0);
    blas.copy(u, w);
    print_vec("blas_copy", w.data(), 5);
    blas.scal(2.0, w);
    print_vec("blas_scal", w.data(), 5);
    std::printf("blas_dot %a\n", blas.dot(u, v));
    std::printf("blas_nrm2 %a\n", blas.nrm2(std::vector<double>{3.0, 4.0, 12.0}));
    blas.axpy(-0.5, u, v);
    print_vec("blas_axpy", v.data(), 5);
    voxel_context();
    return 0;
}
