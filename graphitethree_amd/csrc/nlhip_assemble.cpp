// C ABI of the device least-squares assembly (include/nl_hip.h) and the
// read-back of a resident system's CRS.  The sums are those of a
// least-squares OpenNL context fed the same rows (k_assemble.hip).
#include <algorithm>
#include <cmath>
#include <memory>

#include "nlhip_internal.h"

namespace nlhip {

namespace {

// A device copy of a host array, or the caller's device pointer as is.
template <typename T>
const T* on_device(const T* p, size_t n, DevBuf<T>& keep, hipStream_t s) {
    if (!p || n == 0) return p;
    if (is_device_pointer(p)) return p;
    keep.alloc(n);
    copy_in(keep.get(), p, n * sizeof(T), s);
    return keep.get();
}

template <typename T>
void fetch(T* dst, const T* src, size_t n, hipStream_t s) {
    NLHIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDefault, s));
    NLHIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

System* ls_create(uint32_t nvars, uint32_t nrows, const uint32_t* row_ptr, const uint32_t* var,
                  const double* coef, uint32_t nsys, const double* row_rhs,
                  const uint8_t* locked, const double* values, uint32_t flags, double* rhs,
                  uint32_t* nfree_out) {
    if (nvars == 0) throw Error(NLHIP_ERR_INVALID, "no variables");
    if (nsys == 0) throw Error(NLHIP_ERR_INVALID, "nb_systems must be at least 1");
    if (!row_ptr) throw Error(NLHIP_ERR_INVALID, "null row_ptr");
    device();
    Stream st;
    hipStream_t s = st.get();
    uint32_t ends[2] = {0, 0};
    fetch(&ends[0], row_ptr, 1, s);
    fetch(&ends[1], row_ptr + nrows, 1, s);
    if (ends[0] != 0) throw Error(NLHIP_ERR_INVALID, "row_ptr[0] != 0");
    const uint64_t nent = ends[1];
    if (nent && (!var || !coef)) throw Error(NLHIP_ERR_INVALID, "null var/coef");
    // free indices (host: the locks are a byte per variable)
    std::vector<uint8_t> h_locked(nvars, 0);
    if (locked) fetch(h_locked.data(), locked, nvars, s);
    std::vector<uint32_t> h_free(nvars);
    uint32_t nfree = 0;
    for (uint32_t v = 0; v < nvars; ++v) h_free[v] = h_locked[v] ? 0xffffffffu : nfree++;
    if (nfree_out) *nfree_out = nfree;
    if (nfree == 0) throw Error(NLHIP_ERR_INVALID, "every variable is locked");
    DevBuf<uint32_t> d_rp, d_var, d_free;
    DevBuf<double> d_coef, d_rhs, d_values;
    LSInput in;
    in.nrows = nrows;
    in.nvars = nvars;
    in.nsys = nsys;
    in.nfree = nfree;
    in.nent = nent;
    in.rp = on_device(row_ptr, size_t(nrows) + 1, d_rp, s);
    in.var = on_device(var, nent, d_var, s);
    in.coef = on_device(coef, nent, d_coef, s);
    in.row_rhs = on_device(row_rhs, size_t(nsys) * nrows, d_rhs, s);
    in.values = locked ? on_device(values, size_t(nsys) * nvars, d_values, s) : nullptr;
    d_free.alloc(nvars);
    copy_in(d_free.get(), h_free.data(), size_t(nvars) * 4, s);
    in.free_index = d_free.get();
    LSOutput out;
    ls_assemble_device(in, out, s);
    if (rhs) {
        NLHIP_CHECK(hipMemcpyAsync(rhs, out.rhs.get(), size_t(nsys) * nfree * 8,
                                   hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
    }
    return system_adopt(nfree, out.nnz, std::move(out.rowptr), std::move(out.colind),
                        std::move(out.val), flags);
}

System* lscm_create(uint32_t nv, const double* points, uint32_t nt, const uint32_t* triangles,
                    const uint32_t pin[2], const double uv[4], uint32_t flags, double* rhs,
                    uint32_t* nfree_out) {
    if (nv < 3 || nt == 0 || !points || !triangles)
        throw Error(NLHIP_ERR_INVALID, "empty mesh");
    if (pin[0] == pin[1] || pin[0] >= nv || pin[1] >= nv)
        throw Error(NLHIP_ERR_INVALID, "pins must be two distinct vertices");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(uv[k])) throw Error(NLHIP_ERR_NONFINITE, "pin position is NaN or Inf");
    if (uint64_t(nv) * 2 > 0xffffffffull) throw Error(NLHIP_ERR_UNSUPPORTED, "too many vertices");
    device();
    Stream st;
    hipStream_t s = st.get();
    DevBuf<double> d_p;
    DevBuf<uint32_t> d_t;
    const double* P = on_device(points, 3 * size_t(nv), d_p, s);
    const uint32_t* T = on_device(triangles, 3 * size_t(nt), d_t, s);
    LSRows rows;
    lscm_rows_device(nv, P, nt, T, rows, s);
    const uint32_t nvars = 2 * nv;
    std::vector<uint32_t> h_free(nvars);
    std::vector<double> h_values(nvars, 0.0);
    uint32_t nfree = 0;
    for (uint32_t v = 0; v < nvars; ++v) {
        const bool lk = (v / 2 == pin[0]) || (v / 2 == pin[1]);
        h_free[v] = lk ? 0xffffffffu : nfree++;
    }
    h_values[2 * pin[0]] = uv[0];
    h_values[2 * pin[0] + 1] = uv[1];
    h_values[2 * pin[1]] = uv[2];
    h_values[2 * pin[1] + 1] = uv[3];
    if (nfree_out) *nfree_out = nfree;
    DevBuf<uint32_t> d_free(nvars);
    DevBuf<double> d_values(nvars);
    copy_in(d_free.get(), h_free.data(), size_t(nvars) * 4, s);
    copy_in(d_values.get(), h_values.data(), size_t(nvars) * 8, s);
    LSInput in;
    in.nrows = rows.nrows;
    in.nvars = nvars;
    in.nsys = 1;
    in.nfree = nfree;
    in.nent = 6ull * rows.nrows;
    in.rp = rows.rp.get();
    in.var = rows.var.get();
    in.coef = rows.coef.get();
    in.free_index = d_free.get();
    in.values = d_values.get();
    LSOutput out;
    ls_assemble_device(in, out, s);
    if (rhs) {
        NLHIP_CHECK(hipMemcpyAsync(rhs, out.rhs.get(), size_t(nfree) * 8, hipMemcpyDefault, s));
        NLHIP_CHECK(hipStreamSynchronize(s));
    }
    return system_adopt(nfree, out.nnz, std::move(out.rowptr), std::move(out.colind),
                        std::move(out.val), flags);
}

}  // namespace nlhip

using namespace nlhip;

static System* as_system(nlhip_system h) {
    if (!h) throw Error(NLHIP_ERR_INVALID, "null system handle");
    return reinterpret_cast<System*>(h);
}

extern "C" {

NLAPI int NLAPIENTRY nlhipLeastSquaresCreate(uint32_t nb_variables, uint32_t nb_rows,
                                             const uint32_t* row_ptr, const uint32_t* var,
                                             const double* coef, uint32_t nb_systems,
                                             const double* row_rhs, const uint8_t* locked,
                                             const double* values, uint32_t flags,
                                             nlhip_system* out, double* rhs,
                                             uint32_t* nb_free) {
    return guarded("nlhipLeastSquaresCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        System* S = ls_create(nb_variables, nb_rows, row_ptr, var, coef, nb_systems, row_rhs,
                              locked, values, flags, rhs, nb_free);
        *out = reinterpret_cast<nlhip_system>(S);
    });
}

NLAPI int NLAPIENTRY nlhipLSCMCreate(uint32_t nb_vertices, const double* points,
                                     uint32_t nb_triangles, const uint32_t* triangles,
                                     uint32_t pin0, double u0, double v0, uint32_t pin1,
                                     double u1, double v1, uint32_t flags, nlhip_system* out,
                                     double* rhs, uint32_t* nb_free) {
    return guarded("nlhipLSCMCreate", [&] {
        if (!out) throw Error(NLHIP_ERR_INVALID, "null output handle");
        *out = nullptr;
        const uint32_t pin[2] = {pin0, pin1};
        const double uv[4] = {u0, v0, u1, v1};
        System* S = lscm_create(nb_vertices, points, nb_triangles, triangles, pin, uv, flags,
                                rhs, nb_free);
        *out = reinterpret_cast<nlhip_system>(S);
    });
}

NLAPI int NLAPIENTRY nlhipSystemGetCRS(nlhip_system s, uint32_t* rowptr, uint32_t* colind,
                                       double* val) {
    return guarded("nlhipSystemGetCRS", [&] {
        System* S = as_system(s);
        std::lock_guard<std::mutex> lock(S->mu);
        hipStream_t st = S->stream.get();
        if (rowptr) copy_out(rowptr, S->rowptr.get(), (size_t(S->n) + 1) * 4, st);
        if (colind && S->nnz) copy_out(colind, S->colind.get(), S->nnz * 4, st);
        if (val && S->nnz) copy_out(val, S->val.get(), S->nnz * 8, st);
        S->stream.sync();
    });
}

}  // extern "C"
