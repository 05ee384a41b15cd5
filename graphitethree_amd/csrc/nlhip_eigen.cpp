// Generalized symmetric eigenpairs A x = lambda M x (M a positive diagonal,
// the lumped mass; identity when absent) nearest above a shift sigma: block
// shift-invert subspace iteration with Rayleigh-Ritz.  This replaces the
// ARPACK shift-invert solves that geogram runs for spectral LSCM and
// manifold harmonics (src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:761-763, mesh_grob_spectral_commands.cpp:
// 74-116); the factorized solve ARPACK needs becomes a block Jacobi-PCG
// (nlhipSolveMulti's kernels) on A - sigma M.
//
// Per outer iteration, for a block X of P M-orthonormal vectors:
//   Y  = (A - sigma M)^-1 M X     (block PCG, warm-started from X / (theta - sigma),
//                                   to a tolerance that follows the outer residual)
//   H  = Y^T A Y,  G = Y^T M Y     (fixed-order device reductions)
//   H c = theta G c                (host, P x P: Cholesky of the scaled G,
//                                   cyclic Jacobi on L^-1 H L^-T)
//   X  = Y C, AX = (A Y) C; residuals |AX - theta M X| / (max(|theta|, |sigma|) |MX|)
// until the k wanted pairs (k < P) are below eps.  P is 8, 16 or 32: the
// extra vectors speed convergence of the wanted ones.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>

#include "nlhip_internal.h"

namespace nlhip {

void eig_shift_diag(uint32_t n, const uint32_t* rp, const uint32_t* ci, double* val,
                    const double* mass, double sigma, hipStream_t s);
void eig_scale_rows(uint32_t n, uint32_t p, const double* mass, const double* X, double* Y,
                    hipStream_t s);
void eig_gram(uint32_t n, uint32_t p, const double* Y, const double* Z, double* G,
              DevBuf<double>& part, hipStream_t s);
void eig_combine(uint32_t n, uint32_t p, const double* Y, const double* C, double* X,
                 hipStream_t s);
void eig_residual(uint32_t n, uint32_t p, const double* AX, const double* X,
                  const double* mass, const double* theta, double* out,
                  DevBuf<double>& part, hipStream_t s);

namespace {

// Symmetric P x P (row-major) eigen-decomposition by cyclic Jacobi
// rotations: A = Q diag(w) Q^T.
void jacobi_eigen(int p, std::vector<double> A, std::vector<double>& w, std::vector<double>& Q) {
    Q.assign(size_t(p) * p, 0.0);
    for (int i = 0; i < p; ++i) Q[size_t(i) * p + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < p; ++j) {
                const double v = A[size_t(i) * p + j] * A[size_t(i) * p + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int a = 0; a < p - 1; ++a)
            for (int b = a + 1; b < p; ++b) {
                const double apq = A[size_t(a) * p + b];
                if (apq == 0.0) continue;
                const double app = A[size_t(a) * p + a], aqq = A[size_t(b) * p + b];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < p; ++k) {  // A = J^T A J
                    const double akp = A[size_t(k) * p + a], akq = A[size_t(k) * p + b];
                    A[size_t(k) * p + a] = c * akp - sn * akq;
                    A[size_t(k) * p + b] = sn * akp + c * akq;
                }
                for (int k = 0; k < p; ++k) {
                    const double apk = A[size_t(a) * p + k], aqk = A[size_t(b) * p + k];
                    A[size_t(a) * p + k] = c * apk - sn * aqk;
                    A[size_t(b) * p + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < p; ++k) {
                    const double qkp = Q[size_t(k) * p + a], qkq = Q[size_t(k) * p + b];
                    Q[size_t(k) * p + a] = c * qkp - sn * qkq;
                    Q[size_t(k) * p + b] = sn * qkp + c * qkq;
                }
            }
    }
    w.resize(p);
    for (int i = 0; i < p; ++i) w[i] = A[size_t(i) * p + i];
}

// H c = theta G c for symmetric H, G (G positive definite), ascending theta;
// C (row-major, column j = eigenvector j) with C^T G C = I.  Columns of the
// basis are scaled to unit G-norm first.  Returns false when G is not
// numerically positive definite.
bool rayleigh_ritz(int p, std::vector<double> H, std::vector<double> G, std::vector<double>& theta,
                   std::vector<double>& C) {
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < i; ++j) {
            const double h = 0.5 * (H[size_t(i) * p + j] + H[size_t(j) * p + i]);
            const double g = 0.5 * (G[size_t(i) * p + j] + G[size_t(j) * p + i]);
            H[size_t(i) * p + j] = H[size_t(j) * p + i] = h;
            G[size_t(i) * p + j] = G[size_t(j) * p + i] = g;
        }
    std::vector<double> d(p);
    for (int i = 0; i < p; ++i) {
        if (!(G[size_t(i) * p + i] > 0.0)) return false;
        d[i] = 1.0 / std::sqrt(G[size_t(i) * p + i]);
    }
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            H[size_t(i) * p + j] *= d[i] * d[j];
            G[size_t(i) * p + j] *= d[i] * d[j];
        }
    // G = L L^T
    std::vector<double> L(size_t(p) * p, 0.0);
    for (int j = 0; j < p; ++j) {
        double s = G[size_t(j) * p + j];
        for (int k = 0; k < j; ++k) s -= L[size_t(j) * p + k] * L[size_t(j) * p + k];
        if (!(s > 1e-14)) return false;
        L[size_t(j) * p + j] = std::sqrt(s);
        for (int i = j + 1; i < p; ++i) {
            double t = G[size_t(i) * p + j];
            for (int k = 0; k < j; ++k) t -= L[size_t(i) * p + k] * L[size_t(j) * p + k];
            L[size_t(i) * p + j] = t / L[size_t(j) * p + j];
        }
    }
    // K = L^-1 H L^-T: W = L^-1 H, K = W L^-T = (L^-1 W^T)^T
    auto lower_solve = [&](std::vector<double>& B) {  // B = L^-1 B (columns)
        for (int c = 0; c < p; ++c)
            for (int i = 0; i < p; ++i) {
                double t = B[size_t(i) * p + c];
                for (int k = 0; k < i; ++k) t -= L[size_t(i) * p + k] * B[size_t(k) * p + c];
                B[size_t(i) * p + c] = t / L[size_t(i) * p + i];
            }
    };
    std::vector<double> W = H;
    lower_solve(W);
    std::vector<double> WT(size_t(p) * p);
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) WT[size_t(i) * p + j] = W[size_t(j) * p + i];
    lower_solve(WT);  // WT = L^-1 W^T = (W L^-T)^T = K^T = K
    std::vector<double> w, Q;
    jacobi_eigen(p, WT, w, Q);
    std::vector<int> order(p);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return w[a] < w[b]; });
    // C = D L^-T Q (back substitution with L^T), columns in ascending order
    theta.resize(p);
    C.assign(size_t(p) * p, 0.0);
    for (int jj = 0; jj < p; ++jj) {
        const int j = order[jj];
        theta[jj] = w[j];
        std::vector<double> y(p);
        for (int i = p - 1; i >= 0; --i) {
            double t = Q[size_t(i) * p + j];
            for (int k = i + 1; k < p; ++k) t -= L[size_t(k) * p + i] * y[k];
            y[i] = t / L[size_t(i) * p + i];
        }
        for (int i = 0; i < p; ++i) C[size_t(i) * p + jj] = d[i] * y[i];
    }
    return true;
}

// Deterministic start block (splitmix64 uniforms in [-1, 1)).
void start_block(uint32_t n, uint32_t p, std::vector<double>& X) {
    X.resize(size_t(n) * p);
    uint64_t s = 0x6a09e667f3bcc909ull;
    for (double& v : X) {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v = 2.0 * (double(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

}  // namespace

void eigen_solve(System& A, const double* mass_in, uint32_t k, double sigma, double eps,
                 uint32_t max_iter, double* evals, double* evecs, nlhip_eigen_stats& st) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A.n;
    if (k == 0 || k > 24) throw Error(NLHIP_ERR_INVALID, "nb_eigens must be in [1, 24]");
    if (k >= n) throw Error(NLHIP_ERR_INVALID, "nb_eigens must be below the system size");
    if (!std::isfinite(sigma) || !(eps > 0.0)) throw Error(NLHIP_ERR_INVALID, "bad shift or eps");
    uint32_t P = k <= 4 ? 8 : (k <= 10 ? 16 : 32);
    if (P > n) P = n >= 8 ? 8 : n;
    if (P != 8 && P != 16 && P != 32) throw Error(NLHIP_ERR_INVALID, "system too small");
    Stream stream;
    hipStream_t s = stream.get();
    // mass on the device (checked positive and finite)
    DevBuf<double> d_mass;
    const double* mass = nullptr;
    if (mass_in) {
        std::vector<double> hm(n);
        NLHIP_CHECK(hipMemcpyAsync(hm.data(), mass_in, size_t(n) * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        for (double m : hm)
            if (!(m > 0.0) || !std::isfinite(m))
                throw Error(NLHIP_ERR_INVALID, "mass entries must be positive and finite");
        d_mass.alloc(n);
        copy_in(d_mass.get(), hm.data(), size_t(n) * 8, s);
        mass = d_mass.get();
    }
    // B = A - sigma M, a resident system of its own (Jacobi on its diagonal)
    std::unique_ptr<System> B;
    {
        std::lock_guard<std::mutex> lock(A.mu);
        DevBuf<uint32_t> rp(size_t(n) + 1), ci(A.nnz);
        DevBuf<double> va(A.nnz);
        hipStream_t as = A.stream.get();
        NLHIP_CHECK(hipMemcpyAsync(rp.get(), A.rowptr.get(), (size_t(n) + 1) * 4,
                                   hipMemcpyDeviceToDevice, as));
        NLHIP_CHECK(hipMemcpyAsync(ci.get(), A.colind.get(), A.nnz * 4, hipMemcpyDeviceToDevice, as));
        NLHIP_CHECK(hipMemcpyAsync(va.get(), A.val.get(), A.nnz * 8, hipMemcpyDeviceToDevice, as));
        A.stream.sync();
        eig_shift_diag(n, rp.get(), ci.get(), va.get(), mass, sigma, s);
        B.reset(system_adopt(n, A.nnz, std::move(rp), std::move(ci), std::move(va), 0));
    }
    const size_t blk = size_t(n) * P;
    DevBuf<double> X(blk), Y(blk), MX(blk), AY(blk), AX(blk), part, dG(size_t(P) * P),
        dH(size_t(P) * P), dC(size_t(P) * P), dth(P), dres(3 * size_t(P));
    {
        std::vector<double> h;
        start_block(n, P, h);
        copy_in(X.get(), h.data(), blk * 8, s);
    }
    std::vector<double> theta(P, 0.0), C, G(size_t(P) * P), H(size_t(P) * P);
    // Inner solves are inexact: their relative residual follows the outer
    // one (a hundredth of the largest residual of the wanted pairs, from
    // 1e-2 down to eps / 100), which is all subspace iteration needs and
    // saves most of the block-PCG iterations on ill-conditioned A - sigma M.
    nlhip_solve_params ip;
    nlhipSolveParamsDefault(&ip);
    ip.max_iter = 20000;
    const double inner_floor = std::max(1e-14, eps * 1e-2);
    double outer_res = 1.0;
    st = nlhip_eigen_stats{};
    bool have_theta = false;
    std::vector<double> res(3 * size_t(P));
    const SpmvOps<double> o = A.ops<double>();
    for (uint32_t it = 0; it < max_iter; ++it) {
        // Y = (A - sigma M)^-1 M X, from X / (theta - sigma)
        eig_scale_rows(n, P, mass, X.get(), MX.get(), s);
        if (have_theta) {
            std::vector<double> Cg(size_t(P) * P, 0.0);
            for (uint32_t j = 0; j < P; ++j) {
                const double gap = theta[j] - sigma;
                Cg[size_t(j) * P + j] = gap > 0.0 ? 1.0 / gap : 0.0;
            }
            copy_in(dC.get(), Cg.data(), Cg.size() * 8, s);
            eig_combine(n, P, X.get(), dC.get(), Y.get(), s);
        } else {
            NLHIP_CHECK(hipMemsetAsync(Y.get(), 0, blk * 8, s));
        }
        NLHIP_CHECK(hipStreamSynchronize(s));
        ip.eps = std::max(inner_floor, std::min(1e-2, 1e-2 * outer_res));
        nlhip_solve_stats tot{};
        solve_multi(*B, P, MX.get(), Y.get(), ip, nullptr, &tot);
        st.inner_iterations += tot.iterations;
        // A Y, M Y, Gram matrices
        for (uint32_t j = 0; j < P; ++j)
            NLHIP_CHECK(spmv<double>(n, o.rp, o.cols, o.val, Y.get() + size_t(j) * n,
                                     AY.get() + size_t(j) * n, o.code, A.grid, s));
        eig_scale_rows(n, P, mass, Y.get(), MX.get(), s);
        eig_gram(n, P, Y.get(), AY.get(), dH.get(), part, s);
        eig_gram(n, P, Y.get(), MX.get(), dG.get(), part, s);
        copy_out(H.data(), dH.get(), H.size() * 8, s);
        copy_out(G.data(), dG.get(), G.size() * 8, s);
        NLHIP_CHECK(hipStreamSynchronize(s));
        if (!rayleigh_ritz(int(P), H, G, theta, C))
            throw Error(NLHIP_ERR_HIP, "eigensolver: the iterated block lost rank");
        have_theta = true;
        copy_in(dC.get(), C.data(), C.size() * 8, s);
        eig_combine(n, P, Y.get(), dC.get(), X.get(), s);
        eig_combine(n, P, AY.get(), dC.get(), AX.get(), s);
        copy_in(dth.get(), theta.data(), P * 8, s);
        eig_residual(n, P, AX.get(), X.get(), mass, dth.get(), dres.get(), part, s);
        copy_out(res.data(), dres.get(), res.size() * 8, s);
        NLHIP_CHECK(hipStreamSynchronize(s));
        st.outer_iterations = it + 1;
        st.converged = 0;
        st.max_rel_residual = 0.0;
        // |A x - theta M x| / (max(|theta|, |sigma|) |M x|): relative to the
        // eigenvalue, or to the shift for eigenvalues near 0 (the constant
        // vector of a Laplacian)
        const double big = std::fabs(theta[P - 1]);
        for (uint32_t j = 0; j < k; ++j) {
            double scale = std::max(std::fabs(theta[j]), std::fabs(sigma));
            if (scale == 0.0) scale = big;
            const double den = scale * std::sqrt(res[2 * P + j]);
            const double r = den > 0.0 ? std::sqrt(res[j]) / den : 0.0;
            st.max_rel_residual = std::max(st.max_rel_residual, r);
            if (r <= eps) ++st.converged;
        }
        outer_res = st.max_rel_residual;
        if (st.converged == k) break;
    }
    for (uint32_t j = 0; j < k; ++j) evals[j] = theta[j];
    if (evecs) {
        NLHIP_CHECK(hipMemcpyAsync(evecs, X.get(), size_t(n) * k * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
    }
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace nlhip

using namespace nlhip;

extern "C" {

NLAPI int NLAPIENTRY nlhipEigenSolve(nlhip_system a, const double* mass, uint32_t nb_eigens,
                                     double shift, double eps, uint32_t max_iter,
                                     double* eigenvalues, double* eigenvectors,
                                     nlhip_eigen_stats* stats) {
    return guarded("nlhipEigenSolve", [&] {
        if (!a) throw Error(NLHIP_ERR_INVALID, "null system handle");
        if (!eigenvalues) throw Error(NLHIP_ERR_INVALID, "null eigenvalue output");
        nlhip_eigen_stats st{};
        eigen_solve(*reinterpret_cast<System*>(a), mass, nb_eigens, shift, eps,
                    max_iter ? max_iter : 1000, eigenvalues, eigenvectors, st);
        if (stats) *stats = st;
    });
}

}  // extern "C"
