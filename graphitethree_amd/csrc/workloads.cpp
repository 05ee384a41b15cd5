// Synthetic workloads for the OpenNL solve path: the systems that geogram's
// mesh_compute_LSCM (called at src/lib/OGF/mesh/commands/
// mesh_grob_surface_commands.cpp:759), mesh_make_atlas (:724) and mesh_smooth
// (:703) hand to the solver, rebuilt directly in CRS form.
//
//  * Meshes: a k x k height-field grid over [0,1]^2 (each quad split into two
//    triangles) and disk-like charts cut from such grids.
//  * LSCM (Levy, Petitjean, Ray, Maillot, SIGGRAPH 2002): per triangle the
//    complex conformality row sum_j W_j U_j / sqrt(2 area) = 0, split into a
//    real and an imaginary row, assembled as least-squares normal equations
//    with the two pinned vertices moved to the right-hand side.  The values
//    are accumulated in exactly the order a row-by-row least-squares context
//    would (triangle order, real row before imaginary row), so the result is
//    bitwise equal to feeding wl_lscm_rows() through nlBegin/nlEnd(NL_ROW).
//  * Cotangent Laplacian (Pinkall & Polthier 1993): A = I + t L.
//
// Host-only C++ (OpenMP); no GPU code.  Deterministic from a 64-bit seed via
// splitmix64.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline double uniform01(uint64_t& s) {
    return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

inline double uniform(uint64_t& s, double a, double b) {
    return a + (b - a) * uniform01(s);
}

struct Waves {
    double a[4], f[4], th[4], ph[4];
    void draw(uint64_t& s) {
        for (int m = 0; m < 4; ++m) {
            a[m] = uniform(s, 0.0, 0.1);
            f[m] = uniform(s, 1.0, 8.0);
            th[m] = uniform(s, 0.0, M_PI);
            ph[m] = uniform(s, 0.0, 2.0 * M_PI);
        }
    }
    double z(double x, double y) const {
        double r = 0.0;
        for (int m = 0; m < 4; ++m) {
            r += a[m] * std::sin(2.0 * M_PI * f[m] *
                                     (x * std::cos(th[m]) + y * std::sin(th[m])) +
                                 ph[m]);
        }
        return r;
    }
};

struct Mesh {
    std::vector<double> P;    // 3 * nv
    std::vector<uint32_t> T;  // 3 * nt
    uint32_t nv() const { return uint32_t(P.size() / 3); }
    uint32_t nt() const { return uint32_t(T.size() / 3); }
};

struct System {
    uint32_t n = 0;
    uint32_t nrhs = 0;
    std::vector<uint32_t> rowptr, colind;
    std::vector<double> val;
    std::vector<double> rhs;        // nrhs blocks of n (column-major)
    std::vector<uint32_t> var_of;   // free index -> variable id (LSCM)
};

// k x k grid, vertex j*k+i at (i/(k-1), j/(k-1), z); quad (i,j) split into
// (v00, v10, v11) and (v00, v11, v01).
Mesh* grid_mesh(uint32_t k, uint64_t seed, bool flat) {
    Mesh* M = new Mesh;
    uint64_t s = seed;
    Waves w;
    w.draw(s);
    if (flat) {
        for (double& a : w.a) a = 0.0;
    }
    M->P.resize(size_t(k) * k * 3);
    const double h = 1.0 / double(k - 1);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < long(k); ++j) {
        for (uint32_t i = 0; i < k; ++i) {
            size_t v = size_t(j) * k + i;
            double x = double(i) * h, y = double(j) * h;
            M->P[3 * v] = x;
            M->P[3 * v + 1] = y;
            M->P[3 * v + 2] = flat ? 0.0 : w.z(x, y);
        }
    }
    M->T.resize(size_t(k - 1) * (k - 1) * 6);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < long(k - 1); ++j) {
        for (uint32_t i = 0; i + 1 < k; ++i) {
            uint32_t v00 = uint32_t(j) * k + i, v10 = v00 + 1;
            uint32_t v01 = v00 + k, v11 = v01 + 1;
            size_t t = (size_t(j) * (k - 1) + i) * 6;
            uint32_t* q = &M->T[t];
            q[0] = v00; q[1] = v10; q[2] = v11;
            q[3] = v00; q[4] = v11; q[5] = v01;
        }
    }
    return M;
}

// Disk-like chart with about N vertices: the N grid vertices nearest the
// centre of a g x g grid (ties broken by index), the triangles whose three
// corners are all kept, restricted to the edge-connected component holding
// the centre, re-indexed in grid order.
Mesh* chart_mesh(uint32_t N, uint64_t seed) {
    if (N < 3) N = 3;
    uint32_t g = uint32_t(std::ceil(std::sqrt(4.0 * double(N) / M_PI))) + 3;
    Mesh* G = grid_mesh(g, seed, false);
    const double c = 0.5 * double(g - 1);
    std::vector<uint32_t> order(size_t(g) * g);
    std::iota(order.begin(), order.end(), 0u);
    auto d2 = [&](uint32_t v) {
        double di = double(v % g) - c, dj = double(v / g) - c;
        return di * di + dj * dj;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return d2(a) < d2(b); });
    std::vector<uint8_t> keep(size_t(g) * g, 0);
    for (uint32_t q = 0; q < N && q < order.size(); ++q) keep[order[q]] = 1;
    std::vector<uint32_t> tris;
    for (uint32_t t = 0; t < G->nt(); ++t) {
        const uint32_t* q = &G->T[3 * size_t(t)];
        if (keep[q[0]] && keep[q[1]] && keep[q[2]]) {
            tris.insert(tris.end(), q, q + 3);
        }
    }
    // connected component (through shared edges) of the first kept triangle
    // touching the centre-most vertex
    uint32_t nt = uint32_t(tris.size() / 3);
    std::vector<std::vector<uint32_t>> vt(size_t(g) * g);
    for (uint32_t t = 0; t < nt; ++t)
        for (int c3 = 0; c3 < 3; ++c3) vt[tris[3 * t + c3]].push_back(t);
    std::vector<uint8_t> comp(nt, 0);
    std::vector<uint32_t> stack;
    uint32_t centre = order[0];
    if (!vt[centre].empty()) {
        comp[vt[centre][0]] = 1;
        stack.push_back(vt[centre][0]);
    }
    while (!stack.empty()) {
        uint32_t t = stack.back();
        stack.pop_back();
        for (int e = 0; e < 3; ++e) {
            uint32_t a = tris[3 * t + e], b = tris[3 * t + (e + 1) % 3];
            for (uint32_t u : vt[a]) {
                if (comp[u]) continue;
                const uint32_t* q = &tris[3 * size_t(u)];
                if (q[0] == b || q[1] == b || q[2] == b) {
                    comp[u] = 1;
                    stack.push_back(u);
                }
            }
        }
    }
    std::vector<uint32_t> newid(size_t(g) * g, UINT32_MAX);
    std::vector<uint8_t> used(size_t(g) * g, 0);
    for (uint32_t t = 0; t < nt; ++t)
        if (comp[t])
            for (int c3 = 0; c3 < 3; ++c3) used[tris[3 * t + c3]] = 1;
    Mesh* M = new Mesh;
    uint32_t nv = 0;
    for (uint32_t v = 0; v < g * g; ++v) {
        if (!used[v]) continue;
        newid[v] = nv++;
        M->P.insert(M->P.end(), &G->P[3 * size_t(v)], &G->P[3 * size_t(v)] + 3);
    }
    for (uint32_t t = 0; t < nt; ++t) {
        if (!comp[t]) continue;
        for (int c3 = 0; c3 < 3; ++c3) M->T.push_back(newid[tris[3 * t + c3]]);
    }
    delete G;
    return M;
}

// The 12 LSCM coefficients of triangle t: real row then imaginary row, each
// in the order (u0, v0, u1, v1, u2, v2).  Returns false for a degenerate
// triangle (zero area).
bool lscm_coeffs(const Mesh& M, uint32_t t, double c[12]) {
    const uint32_t* q = &M.T[3 * size_t(t)];
    const double* p0 = &M.P[3 * size_t(q[0])];
    const double* p1 = &M.P[3 * size_t(q[1])];
    const double* p2 = &M.P[3 * size_t(q[2])];
    double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                    e1[0] * e2[1] - e1[1] * e2[0]};
    double dT = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    if (!(dT > 0.0) || !(l1 > 0.0)) return false;
    double X[3] = {e1[0] / l1, e1[1] / l1, e1[2] / l1};
    double Nn[3] = {nr[0] / dT, nr[1] / dT, nr[2] / dT};
    double Y[3] = {Nn[1] * X[2] - Nn[2] * X[1], Nn[2] * X[0] - Nn[0] * X[2],
                   Nn[0] * X[1] - Nn[1] * X[0]};
    // local coordinates: z0 = 0, z1 = (l1, 0), z2 = (e2.X, e2.Y)
    double z1x = l1, z1y = 0.0;
    double z2x = e2[0] * X[0] + e2[1] * X[1] + e2[2] * X[2];
    double z2y = e2[0] * Y[0] + e2[1] * Y[1] + e2[2] * Y[2];
    // W0 = z2 - z1, W1 = z0 - z2, W2 = z1 - z0
    double Wr[3] = {z2x - z1x, -z2x, z1x};
    double Wi[3] = {z2y - z1y, -z2y, z1y};
    double sq = std::sqrt(dT);
    for (int j = 0; j < 3; ++j) {
        double cr = Wr[j] / sq, ci = Wi[j] / sq;
        c[2 * j] = cr;           // real row, u_j
        c[2 * j + 1] = -ci;      // real row, v_j
        c[6 + 2 * j] = ci;       // imag row, u_j
        c[6 + 2 * j + 1] = cr;   // imag row, v_j
    }
    return true;
}

struct VertexTriangles {
    std::vector<uint32_t> off, tri;  // CSR: triangles incident to each vertex
};

VertexTriangles incident(const Mesh& M) {
    VertexTriangles I;
    uint32_t nv = M.nv(), nt = M.nt();
    I.off.assign(size_t(nv) + 1, 0);
    for (uint32_t t = 0; t < nt; ++t)
        for (int c = 0; c < 3; ++c) I.off[M.T[3 * size_t(t) + c] + 1]++;
    for (uint32_t v = 0; v < nv; ++v) I.off[v + 1] += I.off[v];
    I.tri.resize(I.off[nv]);
    std::vector<uint32_t> fill(I.off.begin(), I.off.end() - 1);
    for (uint32_t t = 0; t < nt; ++t)
        for (int c = 0; c < 3; ++c) I.tri[fill[M.T[3 * size_t(t) + c]]++] = t;
    return I;  // triangles listed in increasing order per vertex
}

// Sorted neighbours of v (including v) through the triangles in `ok`.
void neighbours(const Mesh& M, const VertexTriangles& I,
                const std::vector<uint8_t>& ok, uint32_t v,
                std::vector<uint32_t>& out) {
    out.clear();
    out.push_back(v);
    for (uint32_t k = I.off[v]; k < I.off[v + 1]; ++k) {
        uint32_t t = I.tri[k];
        if (!ok[t]) continue;
        for (int c = 0; c < 3; ++c) out.push_back(M.T[3 * size_t(t) + c]);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

System* lscm_system(const Mesh& M, const uint32_t pin[2], const double uv[4]) {
    const uint32_t nv = M.nv(), nt = M.nt();
    const uint32_t nvar = 2 * nv;
    std::vector<uint8_t> locked(nvar, 0);
    std::vector<double> value(nvar, 0.0);
    for (int p = 0; p < 2; ++p) {
        locked[2 * pin[p]] = locked[2 * pin[p] + 1] = 1;
        value[2 * pin[p]] = uv[2 * p];
        value[2 * pin[p] + 1] = uv[2 * p + 1];
    }
    std::vector<uint32_t> fidx(nvar, UINT32_MAX);
    System* S = new System;
    for (uint32_t v = 0; v < nvar; ++v) {
        if (!locked[v]) {
            fidx[v] = uint32_t(S->var_of.size());
            S->var_of.push_back(v);
        }
    }
    const uint32_t n = uint32_t(S->var_of.size());
    S->n = n;
    S->nrhs = 1;
    std::vector<uint8_t> ok(nt);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < long(nt); ++t) {
        double c[12];
        ok[t] = lscm_coeffs(M, uint32_t(t), c) ? 1 : 0;
    }
    VertexTriangles I = incident(M);
    // row lengths
    std::vector<uint32_t> rowlen(n, 0);
#pragma omp parallel
    {
        std::vector<uint32_t> nb;
#pragma omp for schedule(dynamic, 1024)
        for (long v = 0; v < long(nv); ++v) {
            if (locked[2 * v]) continue;
            neighbours(M, I, ok, uint32_t(v), nb);
            uint32_t cnt = 0;
            for (uint32_t b : nb) cnt += locked[2 * b] ? 0 : 2;
            rowlen[fidx[2 * v]] = cnt;
            rowlen[fidx[2 * v + 1]] = cnt;
        }
    }
    S->rowptr.assign(size_t(n) + 1, 0);
    uint64_t acc = 0;
    for (uint32_t r = 0; r < n; ++r) {
        S->rowptr[r] = uint32_t(acc);
        acc += rowlen[r];
    }
    S->rowptr[n] = uint32_t(acc);
    S->colind.resize(acc);
    S->val.assign(acc, 0.0);
    S->rhs.assign(n, 0.0);
#pragma omp parallel
    {
        std::vector<uint32_t> nb;
        std::vector<uint8_t> touched;
#pragma omp for schedule(dynamic, 1024)
        for (long va = 0; va < long(nv); ++va) {
            const uint32_t a = uint32_t(va);
            if (locked[2 * a]) continue;
            neighbours(M, I, ok, a, nb);
            // column layout of a's two rows
            for (int ca = 0; ca < 2; ++ca) {
                uint32_t r = fidx[2 * a + ca];
                uint32_t pos = S->rowptr[r];
                for (uint32_t b : nb) {
                    if (locked[2 * b]) continue;
                    S->colind[pos++] = fidx[2 * b];
                    S->colind[pos++] = fidx[2 * b + 1];
                }
            }
            for (uint32_t k = I.off[a]; k < I.off[a + 1]; ++k) {
                uint32_t t = I.tri[k];
                if (!ok[t]) continue;
                double c[12];
                lscm_coeffs(M, t, c);
                const uint32_t* q = &M.T[3 * size_t(t)];
                int ja = (q[0] == a) ? 0 : (q[1] == a) ? 1 : 2;
                for (int row = 0; row < 2; ++row) {
                    const double* cr = c + 6 * row;
                    // right-hand side of this row: 0 - sum locked c x
                    double s = 0.0;
                    bool any_locked = false;
                    for (int e = 0; e < 6; ++e) {
                        uint32_t var = 2 * q[e / 2] + (e & 1);
                        if (locked[var]) {
                            double p = cr[e] * value[var];
                            s = s - p;
                            any_locked = true;
                        }
                    }
                    for (int ca = 0; ca < 2; ++ca) {
                        const double caf = cr[2 * ja + ca];
                        const uint32_t r = fidx[2 * a + ca];
                        const uint32_t rb = S->rowptr[r], re = S->rowptr[r + 1];
                        for (int e = 0; e < 6; ++e) {
                            uint32_t var = 2 * q[e / 2] + (e & 1);
                            if (locked[var]) continue;
                            uint32_t col = fidx[var];
                            uint32_t pos = rb;
                            while (pos < re && S->colind[pos] != col) ++pos;
                            double p = caf * cr[e];
                            S->val[pos] = S->val[pos] + p;
                        }
                        if (any_locked) {
                            double p = caf * s;
                            S->rhs[r] = S->rhs[r] + p;
                        }
                    }
                }
            }
        }
    }
    return S;
}

// A = I + t L with L the cotangent Laplacian (L_ij = -(cot a + cot b)/2 on
// edges, L_ii = -sum_j L_ij).  Entries of each row accumulate over the
// vertex's incident triangles in increasing order.
System* cotan_system(const Mesh& M, double tstep, uint32_t nrhs, int rhs_kind,
                     uint64_t seed, bool identity = true) {
    const uint32_t nv = M.nv(), nt = M.nt();
    System* S = new System;
    S->n = nv;
    S->nrhs = nrhs;
    // half-cotangent of the angle at each corner, stored per (t, corner)
    std::vector<double> hc(size_t(nt) * 3);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < long(nt); ++t) {
        const uint32_t* q = &M.T[3 * size_t(t)];
        for (int k = 0; k < 3; ++k) {
            const double* p = &M.P[3 * size_t(q[k])];
            const double* a = &M.P[3 * size_t(q[(k + 1) % 3])];
            const double* b = &M.P[3 * size_t(q[(k + 2) % 3])];
            double e1[3] = {a[0] - p[0], a[1] - p[1], a[2] - p[2]};
            double e2[3] = {b[0] - p[0], b[1] - p[1], b[2] - p[2]};
            double dt = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
            double cx = e1[1] * e2[2] - e1[2] * e2[1];
            double cy = e1[2] * e2[0] - e1[0] * e2[2];
            double cz = e1[0] * e2[1] - e1[1] * e2[0];
            double cr = std::sqrt(cx * cx + cy * cy + cz * cz);
            hc[3 * size_t(t) + k] = (cr > 0.0) ? 0.5 * dt / cr : 0.0;
        }
    }
    std::vector<uint8_t> ok(nt, 1);
    VertexTriangles I = incident(M);
    std::vector<uint32_t> rowlen(nv);
#pragma omp parallel
    {
        std::vector<uint32_t> nb;
#pragma omp for schedule(dynamic, 1024)
        for (long v = 0; v < long(nv); ++v) {
            neighbours(M, I, ok, uint32_t(v), nb);
            rowlen[v] = uint32_t(nb.size());
        }
    }
    S->rowptr.assign(size_t(nv) + 1, 0);
    uint64_t acc = 0;
    for (uint32_t r = 0; r < nv; ++r) {
        S->rowptr[r] = uint32_t(acc);
        acc += rowlen[r];
    }
    S->rowptr[nv] = uint32_t(acc);
    S->colind.resize(acc);
    S->val.assign(acc, 0.0);
#pragma omp parallel
    {
        std::vector<uint32_t> nb;
#pragma omp for schedule(dynamic, 1024)
        for (long va = 0; va < long(nv); ++va) {
            const uint32_t a = uint32_t(va);
            neighbours(M, I, ok, a, nb);
            const uint32_t rb = S->rowptr[a];
            for (size_t k = 0; k < nb.size(); ++k) S->colind[rb + k] = nb[k];
            double diag = 0.0;
            for (uint32_t k = I.off[a]; k < I.off[a + 1]; ++k) {
                uint32_t t = I.tri[k];
                const uint32_t* q = &M.T[3 * size_t(t)];
                int ja = (q[0] == a) ? 0 : (q[1] == a) ? 1 : 2;
                // edge (a, q[ja+1]) is opposite corner ja+2, and vice versa
                for (int o = 1; o <= 2; ++o) {
                    uint32_t b = q[(ja + o) % 3];
                    double w = hc[3 * size_t(t) + (ja + 3 - o) % 3];
                    uint32_t pos = rb;
                    while (S->colind[pos] != b) ++pos;
                    double p = tstep * w;
                    S->val[pos] = S->val[pos] - p;
                    diag = diag + w;
                }
            }
            uint32_t pos = rb;
            while (S->colind[pos] != a) ++pos;
            S->val[pos] = S->val[pos] + (identity ? 1.0 + tstep * diag : tstep * diag);
        }
    }
    S->rhs.assign(size_t(nv) * nrhs, 0.0);
    if (rhs_kind == 0) {
        for (uint32_t c = 0; c < nrhs; ++c)
            for (uint32_t v = 0; v < nv; ++v)
                S->rhs[size_t(c) * nv + v] = M.P[3 * size_t(v) + (c % 3)];
    } else {
        // standard normal entries (Box-Muller), column by column
        uint64_t s = seed;
        for (size_t i = 0; i < S->rhs.size(); i += 2) {
            double u1 = uniform01(s), u2 = uniform01(s);
            if (u1 < 1e-300) u1 = 1e-300;
            double r = std::sqrt(-2.0 * std::log(u1));
            S->rhs[i] = r * std::cos(2.0 * M_PI * u2);
            if (i + 1 < S->rhs.size()) S->rhs[i + 1] = r * std::sin(2.0 * M_PI * u2);
        }
    }
    return S;
}

// A texture atlas: `count` disk-like charts with N ~ logU[Nmin, Nmax]
// vertices each, their LSCM systems (extreme vertices pinned at their own
// (x, y)) packed block-diagonally.
struct Atlas {
    std::vector<uint32_t> row_off, rowptr, colind, chart_nv;
    std::vector<double> val, rhs;
};

Atlas* build_atlas(uint32_t count, uint64_t seed, double Nmin, double Nmax);

// The meshes of the same atlas, concatenated: chart c owns vertices
// [vertex_off[c], vertex_off[c+1]) and the triangles over them (global ids),
// with its two pins (global ids) and their (u0, v0, u1, v1).
struct AtlasMesh {
    std::vector<uint32_t> vertex_off, tri_off, T, pins;
    std::vector<double> P, pin_uv;
};

AtlasMesh* build_atlas_mesh(uint32_t count, uint64_t seed, double Nmin, double Nmax);

}  // namespace

extern "C" void wl_atlas_chart(uint64_t seed, uint32_t idx, double Nmin, double Nmax,
                               uint32_t* N, uint64_t* chart_seed);
extern "C" void wl_mesh_extreme_pins(const void* mp, uint32_t* p0, uint32_t* p1);

namespace {
Atlas* build_atlas(uint32_t count, uint64_t seed, double Nmin, double Nmax) {
    std::vector<System*> sys(count, nullptr);
    std::vector<uint32_t> nv(count, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < long(count); ++c) {
        uint32_t N;
        uint64_t cs;
        wl_atlas_chart(seed, uint32_t(c), Nmin, Nmax, &N, &cs);
        Mesh* M = chart_mesh(N, cs);
        uint32_t p[2] = {0, 0};
        {
            uint32_t lo = 0, hi = 0;
            for (uint32_t v = 1; v < M->nv(); ++v) {
                double sv = M->P[3 * size_t(v)] + M->P[3 * size_t(v) + 1];
                if (sv < M->P[3 * size_t(lo)] + M->P[3 * size_t(lo) + 1]) lo = v;
                if (sv > M->P[3 * size_t(hi)] + M->P[3 * size_t(hi) + 1]) hi = v;
            }
            p[0] = lo;
            p[1] = hi;
        }
        double uv[4] = {M->P[3 * size_t(p[0])], M->P[3 * size_t(p[0]) + 1],
                        M->P[3 * size_t(p[1])], M->P[3 * size_t(p[1]) + 1]};
        sys[c] = lscm_system(*M, p, uv);
        nv[c] = M->nv();
        delete M;
    }
    Atlas* A = new Atlas;
    A->row_off.push_back(0);
    uint64_t nnz = 0;
    for (uint32_t c = 0; c < count; ++c) {
        A->row_off.push_back(A->row_off.back() + sys[c]->n);
        nnz += sys[c]->colind.size();
    }
    const uint32_t n = A->row_off.back();
    A->rowptr.resize(size_t(n) + 1);
    A->colind.resize(nnz);
    A->val.resize(nnz);
    A->rhs.resize(n);
    A->chart_nv = nv;
    std::vector<uint64_t> nnz_off(count + 1, 0);
    for (uint32_t c = 0; c < count; ++c) nnz_off[c + 1] = nnz_off[c] + sys[c]->colind.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < long(count); ++c) {
        const System& S = *sys[c];
        const uint32_t r0 = A->row_off[c];
        const uint64_t z0 = nnz_off[c];
        for (uint32_t i = 0; i < S.n; ++i) A->rowptr[r0 + i] = uint32_t(z0 + S.rowptr[i]);
        for (size_t j = 0; j < S.colind.size(); ++j) {
            A->colind[z0 + j] = r0 + S.colind[j];
            A->val[z0 + j] = S.val[j];
        }
        std::copy(S.rhs.begin(), S.rhs.begin() + S.n, A->rhs.begin() + r0);
        delete sys[c];
    }
    A->rowptr[n] = uint32_t(nnz);
    return A;
}

AtlasMesh* build_atlas_mesh(uint32_t count, uint64_t seed, double Nmin, double Nmax) {
    std::vector<Mesh*> mesh(count, nullptr);
    std::vector<uint32_t> pin(2 * size_t(count));
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < long(count); ++c) {
        uint32_t N;
        uint64_t cs;
        wl_atlas_chart(seed, uint32_t(c), Nmin, Nmax, &N, &cs);
        mesh[c] = chart_mesh(N, cs);
        wl_mesh_extreme_pins(mesh[c], &pin[2 * c], &pin[2 * c + 1]);
    }
    AtlasMesh* A = new AtlasMesh;
    A->vertex_off.assign(size_t(count) + 1, 0);
    A->tri_off.assign(size_t(count) + 1, 0);
    for (uint32_t c = 0; c < count; ++c) {
        A->vertex_off[c + 1] = A->vertex_off[c] + mesh[c]->nv();
        A->tri_off[c + 1] = A->tri_off[c] + mesh[c]->nt();
    }
    A->P.resize(3 * size_t(A->vertex_off[count]));
    A->T.resize(3 * size_t(A->tri_off[count]));
    A->pins.resize(2 * size_t(count));
    A->pin_uv.resize(4 * size_t(count));
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < long(count); ++c) {
        const Mesh& M = *mesh[c];
        const uint32_t v0 = A->vertex_off[c];
        std::copy(M.P.begin(), M.P.end(), A->P.begin() + 3 * size_t(v0));
        for (size_t j = 0; j < M.T.size(); ++j) A->T[3 * size_t(A->tri_off[c]) + j] = v0 + M.T[j];
        for (int k = 0; k < 2; ++k) {
            const uint32_t p = pin[2 * c + k];
            A->pins[2 * c + k] = v0 + p;
            A->pin_uv[4 * c + 2 * k] = M.P[3 * size_t(p)];
            A->pin_uv[4 * c + 2 * k + 1] = M.P[3 * size_t(p) + 1];
        }
        delete mesh[c];
    }
    return A;
}
}  // namespace

extern "C" {

uint64_t wl_splitmix64(uint64_t* state) { return splitmix64(*state); }
double wl_uniform01(uint64_t* state) { return uniform01(*state); }

void* wl_mesh_grid(uint32_t k, uint64_t seed, int flat) {
    if (k < 2) return nullptr;
    return grid_mesh(k, seed, flat != 0);
}
void* wl_mesh_chart(uint32_t N, uint64_t seed) { return chart_mesh(N, seed); }
uint32_t wl_mesh_nv(const void* m) { return static_cast<const Mesh*>(m)->nv(); }
uint32_t wl_mesh_nt(const void* m) { return static_cast<const Mesh*>(m)->nt(); }
const double* wl_mesh_points(const void* m) {
    return static_cast<const Mesh*>(m)->P.data();
}
const uint32_t* wl_mesh_triangles(const void* m) {
    return static_cast<const Mesh*>(m)->T.data();
}
void wl_mesh_free(void* m) { delete static_cast<Mesh*>(m); }

// Pins: the vertices with the smallest and largest x + y (ties: lowest id).
void wl_mesh_extreme_pins(const void* mp, uint32_t* p0, uint32_t* p1) {
    const Mesh& M = *static_cast<const Mesh*>(mp);
    uint32_t lo = 0, hi = 0;
    for (uint32_t v = 1; v < M.nv(); ++v) {
        double s = M.P[3 * size_t(v)] + M.P[3 * size_t(v) + 1];
        if (s < M.P[3 * size_t(lo)] + M.P[3 * size_t(lo) + 1]) lo = v;
        if (s > M.P[3 * size_t(hi)] + M.P[3 * size_t(hi) + 1]) hi = v;
    }
    *p0 = lo;
    *p1 = hi;
}

// Per triangle: 12 LSCM coefficients (see lscm_coeffs) and a validity flag.
void wl_lscm_rows(const void* mp, double* coeffs, uint8_t* valid) {
    const Mesh& M = *static_cast<const Mesh*>(mp);
    for (uint32_t t = 0; t < M.nt(); ++t) {
        double c[12] = {0};
        valid[t] = lscm_coeffs(M, t, c) ? 1 : 0;
        std::memcpy(coeffs + 12 * size_t(t), c, sizeof(c));
    }
}

// LSCM normal equations with vertex pin0 fixed at (u0,v0) and pin1 at (u1,v1).
void* wl_system_lscm(const void* mp, uint32_t pin0, double u0, double v0,
                     uint32_t pin1, double u1, double v1) {
    const Mesh& M = *static_cast<const Mesh*>(mp);
    if (pin0 == pin1 || pin0 >= M.nv() || pin1 >= M.nv()) return nullptr;
    uint32_t pin[2] = {pin0, pin1};
    double uv[4] = {u0, v0, u1, v1};
    return lscm_system(M, pin, uv);
}

// A = I + t L; rhs_kind 0: columns x, y, z, x, ... of the vertices;
// rhs_kind 1: standard normal entries drawn from `seed`.
// The cotangent stiffness matrix L (no identity term; rows sum to 0) and the
// lumped mass of each vertex (a third of the area of its triangles).
void* wl_system_laplacian(const void* mp, double* mass) {
    const Mesh& M = *static_cast<const Mesh*>(mp);
    for (uint32_t v = 0; v < M.nv(); ++v) mass[v] = 0.0;
    for (uint32_t t = 0; t < M.nt(); ++t) {
        const uint32_t* q = &M.T[3 * size_t(t)];
        const double* p0 = &M.P[3 * size_t(q[0])];
        const double* p1 = &M.P[3 * size_t(q[1])];
        const double* p2 = &M.P[3 * size_t(q[2])];
        double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        double cx = e1[1] * e2[2] - e1[2] * e2[1];
        double cy = e1[2] * e2[0] - e1[0] * e2[2];
        double cz = e1[0] * e2[1] - e1[1] * e2[0];
        double a3 = std::sqrt(cx * cx + cy * cy + cz * cz) / 6.0;
        for (int k = 0; k < 3; ++k) mass[q[k]] += a3;
    }
    return cotan_system(M, 1.0, 1, 0, 0, false);
}

void* wl_system_cotan(const void* mp, double t, uint32_t nrhs, int rhs_kind,
                      uint64_t seed) {
    return cotan_system(*static_cast<const Mesh*>(mp), t, nrhs, rhs_kind, seed);
}

uint32_t wl_sys_n(const void* s) { return static_cast<const System*>(s)->n; }
uint64_t wl_sys_nnz(const void* s) {
    return static_cast<const System*>(s)->colind.size();
}
uint32_t wl_sys_nrhs(const void* s) { return static_cast<const System*>(s)->nrhs; }
const uint32_t* wl_sys_rowptr(const void* s) {
    return static_cast<const System*>(s)->rowptr.data();
}
const uint32_t* wl_sys_colind(const void* s) {
    return static_cast<const System*>(s)->colind.data();
}
const double* wl_sys_val(const void* s) {
    return static_cast<const System*>(s)->val.data();
}
const double* wl_sys_rhs(const void* s) {
    return static_cast<const System*>(s)->rhs.data();
}
const uint32_t* wl_sys_var_of(const void* s) {
    return static_cast<const System*>(s)->var_of.data();
}
void wl_sys_free(void* s) { delete static_cast<System*>(s); }

// Atlas chart parameters: chart `idx` of an atlas drawn from `seed` has
// N = round(exp(U(ln Nmin, ln Nmax))) vertices and its own mesh seed.
void wl_atlas_chart(uint64_t seed, uint32_t idx, double Nmin, double Nmax,
                    uint32_t* N, uint64_t* chart_seed) {
    uint64_t s = seed ^ (0xA5A5A5A5ULL + 0x9E3779B97F4A7C15ULL * (idx + 1));
    splitmix64(s);
    double u = uniform01(s);
    double ln = std::log(Nmin) + u * (std::log(Nmax) - std::log(Nmin));
    *N = uint32_t(std::llround(std::exp(ln)));
    *chart_seed = splitmix64(s);
}

void* wl_atlas(uint32_t count, uint64_t seed, double Nmin, double Nmax) {
    if (count == 0) return nullptr;
    return build_atlas(count, seed, Nmin, Nmax);
}
uint32_t wl_atlas_count(const void* a) {
    return uint32_t(static_cast<const Atlas*>(a)->row_off.size() - 1);
}
uint64_t wl_atlas_nnz(const void* a) { return static_cast<const Atlas*>(a)->colind.size(); }
const uint32_t* wl_atlas_row_off(const void* a) {
    return static_cast<const Atlas*>(a)->row_off.data();
}
const uint32_t* wl_atlas_rowptr(const void* a) {
    return static_cast<const Atlas*>(a)->rowptr.data();
}
const uint32_t* wl_atlas_colind(const void* a) {
    return static_cast<const Atlas*>(a)->colind.data();
}
const double* wl_atlas_val(const void* a) { return static_cast<const Atlas*>(a)->val.data(); }
const double* wl_atlas_rhs(const void* a) { return static_cast<const Atlas*>(a)->rhs.data(); }
const uint32_t* wl_atlas_chart_nv(const void* a) {
    return static_cast<const Atlas*>(a)->chart_nv.data();
}
void wl_atlas_free(void* a) { delete static_cast<Atlas*>(a); }

void* wl_atlas_mesh(uint32_t count, uint64_t seed, double Nmin, double Nmax) {
    if (count == 0) return nullptr;
    return build_atlas_mesh(count, seed, Nmin, Nmax);
}
uint32_t wl_atlas_mesh_nv(const void* a) {
    return static_cast<const AtlasMesh*>(a)->vertex_off.back();
}
uint32_t wl_atlas_mesh_nt(const void* a) {
    return static_cast<const AtlasMesh*>(a)->tri_off.back();
}
const uint32_t* wl_atlas_mesh_vertex_off(const void* a) {
    return static_cast<const AtlasMesh*>(a)->vertex_off.data();
}
const uint32_t* wl_atlas_mesh_tri_off(const void* a) {
    return static_cast<const AtlasMesh*>(a)->tri_off.data();
}
const double* wl_atlas_mesh_points(const void* a) {
    return static_cast<const AtlasMesh*>(a)->P.data();
}
const uint32_t* wl_atlas_mesh_triangles(const void* a) {
    return static_cast<const AtlasMesh*>(a)->T.data();
}
const uint32_t* wl_atlas_mesh_pins(const void* a) {
    return static_cast<const AtlasMesh*>(a)->pins.data();
}
const double* wl_atlas_mesh_pin_uv(const void* a) {
    return static_cast<const AtlasMesh*>(a)->pin_uv.data();
}
void wl_atlas_mesh_free(void* a) { delete static_cast<AtlasMesh*>(a); }

}  // extern "C"
