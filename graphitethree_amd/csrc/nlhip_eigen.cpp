// Generalized symmetric eigenpairs A x = lambda M x (M a positive diagonal,
// the lumped mass; identity when absent): the k smallest, for a shift below
// the spectrum.  This replaces the ARPACK shift-invert solves that geogram
// runs for spectral LSCM and manifold harmonics (src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:761-763, mesh_grob_spectral_commands.cpp:
// 74-116), by Chebyshev-filtered subspace iteration (Zhou and Saad): no
// linear solves, only multi-vector SpMMs.
//
// On a block X of P M-orthonormal Ritz vectors with Ritz values theta:
//   Y  = p_m(M^-1 A) X, p_m the degree-m Chebyshev polynomial that damps
//        [theta_{P-1}, b] (b a Gershgorin bound of the spectrum) and is 1 at
//        the smallest Ritz value; three-term recurrence, one fused SpMM per
//        degree
//   H  = Y^T A Y,  G = Y^T M Y     (fixed-order device reductions)
//   H c = theta G c                (host, P x P: Cholesky of the scaled G,
//                                   cyclic Jacobi on L^-1 H L^-T)
//   X  = Y C, AX = (A Y) C; residuals |AX - theta M X| / (max(|theta|,
//        |shift|) |MX|)
// until the k wanted pairs (k < P) are below eps.  P is 8, 16 or 32: the
// extra vectors keep the damped interval away from the wanted eigenvalues.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>

#include "nlhip_internal.h"

namespace nlhip {

EigShape eig_shape(uint32_t n, uint32_t p, int cu);
void eig_spmm(const EigShape& e, uint32_t n, uint32_t p, const uint32_t* rp, const uint32_t* ci,
              const double* val, const double* Y, double* AY, hipStream_t s);
void eig_cheb(const EigShape& e, uint32_t n, uint32_t p, const uint32_t* rp, const uint32_t* ci,
              const double* val, const double* minv, const double* Y, const double* Xold,
              double* Ynew, double c, double s1, double s2, hipStream_t s);
void eig_gram(uint32_t n, uint32_t p, const double* Y, const double* Z, const double* mass,
              double* G, DevBuf<double>& part, hipStream_t s);
void eig_combine(uint32_t n, uint32_t p, const double* Y, const double* C, double* X,
                 hipStream_t s);
void eig_residual(uint32_t n, uint32_t p, const double* AX, const double* X, const double* mass,
                  const double* theta, double* out, DevBuf<double>& part, hipStream_t s);
double eig_bound(uint32_t n, const uint32_t* rp, const double* val, const double* minv,
                 hipStream_t s);
void eig_inverse(uint32_t n, const double* m, double* minv, hipStream_t s);

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

// Chebyshev degree per outer iteration.
constexpr int kChebDegree = 40;

void eigen_solve(System& A, const double* mass_in, uint32_t k, double sigma, double eps,
                 uint32_t max_iter, double* evals, double* evecs, nlhip_eigen_stats& st) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = A.n;
    if (k == 0 || k > 24) throw Error(NLHIP_ERR_INVALID, "nb_eigens must be in [1, 24]");
    if (!std::isfinite(sigma) || !(eps > 0.0)) throw Error(NLHIP_ERR_INVALID, "bad shift or eps");
    uint32_t P = k <= 4 ? 8 : (k <= 8 ? 16 : 32);
    if (P >= n) throw Error(NLHIP_ERR_INVALID, "system too small for the eigenvector block");
    Stream stream;
    hipStream_t s = stream.get();
    DevBuf<double> d_mass, d_minv;
    const double* mass = nullptr;
    const double* minv = nullptr;
    if (mass_in) {
        std::vector<double> hm(n);
        NLHIP_CHECK(hipMemcpyAsync(hm.data(), mass_in, size_t(n) * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
        for (double m : hm)
            if (!(m > 0.0) || !std::isfinite(m))
                throw Error(NLHIP_ERR_INVALID, "mass entries must be positive and finite");
        d_mass.alloc(n);
        d_minv.alloc(n);
        copy_in(d_mass.get(), hm.data(), size_t(n) * 8, s);
        eig_inverse(n, d_mass.get(), d_minv.get(), s);
        mass = d_mass.get();
        minv = d_minv.get();
    }
    const uint32_t* rp = A.rowptr.get();
    const uint32_t* ci = A.colind.get();
    const double* va = A.val.get();
    EigShape shape = eig_shape(n, P, device().cu_count);
    shape.one_chunk = spmm_one_chunk(n, rp, P, s);
    const double b_up = eig_bound(n, rp, va, minv, s);
    if (!(b_up > 0.0) || !std::isfinite(b_up))
        throw Error(NLHIP_ERR_INVALID, "matrix without a positive spectrum bound");
    const size_t blk = size_t(n) * P;
    DevBuf<double> buf[3] = {DevBuf<double>(blk), DevBuf<double>(blk), DevBuf<double>(blk)};
    DevBuf<double> AY(blk), AX(blk), part, dG(size_t(P) * P), dH(size_t(P) * P),
        dC(size_t(P) * P), dth(P), dres(3 * size_t(P));
    {
        std::vector<double> h;
        start_block(n, P, h);
        copy_in(buf[0].get(), h.data(), blk * 8, s);
    }
    std::vector<double> theta(P, 0.0), C, G(size_t(P) * P), H(size_t(P) * P);
    std::vector<double> res(3 * size_t(P));
    st = nlhip_eigen_stats{};
    int x_at = 0;  // buffer holding the current block X
    // Rayleigh-Ritz on the block in buf[y]: X = Y C into buf[x_to], AX
    auto rayleigh_ritz_step = [&](int y, int x_to) {
        eig_spmm(shape, n, P, rp, ci, va, buf[y].get(), AY.get(), s);
        eig_gram(n, P, buf[y].get(), AY.get(), nullptr, dH.get(), part, s);
        eig_gram(n, P, buf[y].get(), buf[y].get(), mass, dG.get(), part, s);
        copy_out(H.data(), dH.get(), H.size() * 8, s);
        copy_out(G.data(), dG.get(), G.size() * 8, s);
        NLHIP_CHECK(hipStreamSynchronize(s));
        if (!rayleigh_ritz(int(P), H, G, theta, C))
            throw Error(NLHIP_ERR_HIP, "eigensolver: the filtered block lost rank");
        copy_in(dC.get(), C.data(), C.size() * 8, s);
        eig_combine(n, P, buf[y].get(), dC.get(), buf[x_to].get(), s);
        eig_combine(n, P, AY.get(), dC.get(), AX.get(), s);
    };
    rayleigh_ritz_step(0, 1);
    x_at = 1;
    for (uint32_t it = 0; it < max_iter; ++it) {
        // filter: damp [a, b_up], normalised at the smallest Ritz value
        const double a = theta[P - 1];
        const double lo = std::min(theta[0], sigma);
        if (!(b_up > a) || !(a > lo)) throw Error(NLHIP_ERR_HIP, "eigensolver: degenerate filter interval");
        const double e = 0.5 * (b_up - a), c = 0.5 * (b_up + a);
        double sg = e / (lo - c);
        const double sg1 = sg;
        int prev = x_at, cur = (x_at + 1) % 3, nxt = (x_at + 2) % 3;
        eig_cheb(shape, n, P, rp, ci, va, minv, buf[prev].get(), nullptr, buf[cur].get(), c,
                 sg1 / e, 0.0, s);
        for (int d = 2; d <= kChebDegree; ++d) {
            const double sg2 = 1.0 / (2.0 / sg1 - sg);
            eig_cheb(shape, n, P, rp, ci, va, minv, buf[cur].get(), buf[prev].get(),
                     buf[nxt].get(), c, 2.0 * sg2 / e, sg * sg2, s);
            sg = sg2;
            const int t = prev;
            prev = cur;
            cur = nxt;
            nxt = t;
        }
        st.inner_iterations += kChebDegree;
        // Rayleigh-Ritz into a buffer other than the filtered one
        x_at = (cur + 1) % 3;
        rayleigh_ritz_step(cur, x_at);
        copy_in(dth.get(), theta.data(), P * 8, s);
        eig_residual(n, P, AX.get(), buf[x_at].get(), mass, dth.get(), dres.get(), part, s);
        copy_out(res.data(), dres.get(), res.size() * 8, s);
        NLHIP_CHECK(hipStreamSynchronize(s));
        st.outer_iterations = it + 1;
        st.converged = 0;
        st.max_rel_residual = 0.0;
        // |A x - theta M x| / (max(|theta|, |sigma|) |M x|): relative to the
        // eigenvalue, or to the shift for eigenvalues near 0 (the constant
        // vector of a Laplacian).  A pair also counts as converged when its
        // residual is at the rounding level of the operator, 1000 u b |M x|
        // (b the spectrum bound): an eigenvalue near 0 cannot get below it.
        const double big = std::fabs(theta[P - 1]);
        const double floor_rel = 1e3 * 2.220446049250313e-16 * b_up;
        for (uint32_t j = 0; j < k; ++j) {
            double scale = std::max(std::fabs(theta[j]), std::fabs(sigma));
            if (scale == 0.0) scale = big;
            const double mx = std::sqrt(res[2 * P + j]);
            const double rn = std::sqrt(res[j]);
            const double r = scale * mx > 0.0 ? rn / (scale * mx) : 0.0;
            st.max_rel_residual = std::max(st.max_rel_residual, r);
            if (r <= eps || rn <= floor_rel * mx) ++st.converged;
        }
        if (st.converged == k) break;
    }
    for (uint32_t j = 0; j < k; ++j) evals[j] = theta[j];
    if (evecs) {
        // interleaved block -> the first k columns, column-major
        std::vector<double> h(blk), out(size_t(n) * k);
        copy_out(h.data(), buf[x_at].get(), blk * 8, s);
        NLHIP_CHECK(hipStreamSynchronize(s));
        for (uint32_t r = 0; r < n; ++r)
            for (uint32_t j = 0; j < k; ++j) out[size_t(j) * n + r] = h[size_t(r) * P + j];
        NLHIP_CHECK(hipMemcpyAsync(evecs, out.data(), out.size() * 8, hipMemcpyDefault, s));
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
