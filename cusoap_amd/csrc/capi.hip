// extern "C" entry points of libcusoap_amd.so (declared in include/cusoap_amd.h).
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace cusoap {

static thread_local std::string g_last_error;

// ---- event profiler ---------------------------------------------------------
struct ProfSpan {
    hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::map<std::string, std::vector<ProfSpan>> g_prof;
static std::map<std::string, hipEvent_t> g_prof_open;

void prof_begin(const char* name, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    hipEventRecord(e, st);
    // a begin whose end was skipped (an error return in between) leaves its event open:
    // release it instead of leaking it
    auto it = g_prof_open.find(name);
    if (it != g_prof_open.end()) {
        hipEventDestroy(it->second);
        it->second = e;
    } else {
        g_prof_open[name] = e;
    }
}

void prof_end(const char* name, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_on) return;
    auto it = g_prof_open.find(name);
    if (it == g_prof_open.end()) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    hipEventRecord(e, st);
    g_prof[name].push_back({it->second, e});
    g_prof_open.erase(it);
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int nl_plan_impl(const double*, int32_t, int32_t, int32_t, double, void*, cusoap_nl_plan_t*, hipStream_t);
int nl_count_impl(const cusoap_nl_plan_t*, const double*, int32_t, const double*, int32_t, const double*, int32_t,
                  double, int32_t, void*, size_t, int64_t*, int64_t*, int32_t*, hipStream_t);
int nl_fill_impl(const cusoap_nl_plan_t*, const double*, int32_t, const double*, int32_t, const double*, int32_t,
                 double, int32_t, void*, size_t, const int64_t*, int32_t, int32_t*, hipStream_t);
int coefficients_impl(const cusoap_basis_t*, int, const double*, int, const int32_t*, const double*, int,
                      const double*, int, const int32_t*, const int64_t*, const int32_t*, int, void*, hipStream_t);
int power_spectrum_impl(const cusoap_basis_t*, const void*, int, int, const uint32_t*, int, void*, int, hipStream_t);
int cc_projection_impl(const cusoap_basis_t*, const void*, int, int, void*, int, hipStream_t);
int power_spectrum_mfma_impl(const cusoap_basis_t*, const void*, int, int, const int32_t*, int, int, int, void*, int,
                             hipStream_t);
size_t jacobian_ws_bytes(const cusoap_basis_t*, int, int64_t, int, int);
int jacobian_dense_impl(const cusoap_basis_t*, int, const double*, int, const int32_t*, const double*, int,
                        const double*, int, const int32_t*, const int64_t*, const int32_t*, int64_t, const double*,
                        const int32_t*, int, const uint32_t*, int, void*, size_t, void*, int, int, hipStream_t);
size_t vjp_ws_bytes(const cusoap_basis_t*, int);
size_t vjp_ordered_ws_bytes(const cusoap_basis_t*, int, int, int64_t);
size_t edge_reverse_ws_bytes(int);
int edge_reverse_impl(const double*, int, const double*, int, const double*, int, const int64_t*, const int32_t*,
                      void*, size_t, int32_t*, hipStream_t);
int vjp_ordered_impl(const cusoap_basis_t*, int, int, const double*, int, const int32_t*, const double*, int,
                     const double*, int, const int32_t*, const int32_t*, const int64_t*, const int32_t*,
                     const int32_t*, int64_t, const void*, int, const int32_t*, const void*, int, int, int, void*,
                     size_t, double*, hipStream_t);
int vjp_impl(const cusoap_basis_t*, int, int, const double*, int, const int32_t*, const double*, int, const double*,
             int, const int32_t*, const int32_t*, const int64_t*, const int32_t*, const void*, int, const int32_t*,
             const void*, int, int, int, void*, size_t, double*, hipStream_t);
int vjp_noalloc_impl(const cusoap_basis_t*, int, int, const double*, int, const int32_t*, const double*, int,
                     const double*, int, const int32_t*, const int32_t*, const int64_t*, const int32_t*, const void*,
                     int, const int32_t*, const void*, int, int, int, double*, hipStream_t);
int lowdin_impl(const double*, int, int, double, double*, hipStream_t);
int64_t mlp_param_count(int, int);
int radial_exact_impl(const cusoap_basis_t*, const double*, int, double*, double*, hipStream_t);
int mlp_eval_impl(const float*, int, int, int, const float*, int, float*, float*, hipStream_t);
size_t mlp_train_ws_bytes(int, int, int);
int mlp_loss_grad_impl(const float*, int, int, int, const float*, const float*, const float*, int, int64_t,
                       const float*, const float*, int, uint32_t, int64_t, void*, size_t, float*, float*,
                       hipStream_t);
int adam_impl(float*, const float*, float*, float*, int64_t, float, float, float, float, int64_t, hipStream_t);
size_t mlp_refit_ws_bytes(int, int, int);
int mlp_refit_impl(float*, int, int, int, const float*, const float*, const float*, int, int64_t, const float*,
                   double, void*, size_t, hipStream_t);
size_t radial_table_ws_bytes(const cusoap_basis_t*, int);
int radial_table_impl(const cusoap_basis_t*, const float*, int, const int32_t*, const float*, double, int, double,
                      void*, size_t, float*, hipStream_t);

}  // namespace cusoap

using namespace cusoap;

extern "C" {

const char* cusoap_last_error(void) { return g_last_error.c_str(); }
int cusoap_version(void) { return 1; }
int cusoap_max_l(void) { return kMaxL; }
int cusoap_max_n(void) { return kMaxN; }

void cusoap_profile_enable(int32_t on) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
}

int cusoap_profile_get(const char* name, double* total_ms, int32_t* count) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    double tot = 0.0;
    int n = 0;
    auto it = g_prof.find(name);
    if (it != g_prof.end()) {
        // take the spans out of the map first, so that every event is destroyed exactly
        // once whatever happens below
        std::vector<ProfSpan> spans;
        spans.swap(it->second);
        g_prof.erase(it);
        bool sync_ok = true;
        for (auto& sp : spans) {
            float ms = 0.f;
            if (sync_ok && hipEventSynchronize(sp.b) != hipSuccess) sync_ok = false;
            if (sync_ok && hipEventElapsedTime(&ms, sp.a, sp.b) == hipSuccess) {
                tot += ms;
                ++n;
            }
            hipEventDestroy(sp.a);
            hipEventDestroy(sp.b);
        }
        if (!sync_ok) return fail(CUSOAP_ERR_HIP, "profile: event sync failed");
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = n;
    return CUSOAP_OK;
}

int cusoap_nl_plan(const double* centers, int32_t n_centers, int32_t n_atoms, int32_t n_img, double cutoff, void* tmp,
                   cusoap_nl_plan_t* plan, void* stream) {
    return nl_plan_impl(centers, n_centers, n_atoms, n_img, cutoff, tmp, plan, (hipStream_t)stream);
}

int cusoap_nl_count(const cusoap_nl_plan_t* plan, const double* positions, int32_t n_atoms, const double* shifts,
                    int32_t n_img, const double* centers, int32_t n_centers, double cutoff, int32_t drop_self,
                    void* ws, size_t ws_bytes, int64_t* offsets, int64_t* total_host, int32_t* max_count_host,
                    void* stream) {
    return nl_count_impl(plan, positions, n_atoms, shifts, n_img, centers, n_centers, cutoff, drop_self, ws, ws_bytes,
                         offsets, total_host, max_count_host, (hipStream_t)stream);
}

int cusoap_nl_fill(const cusoap_nl_plan_t* plan, const double* positions, int32_t n_atoms, const double* shifts,
                   int32_t n_img, const double* centers, int32_t n_centers, double cutoff, int32_t drop_self, void* ws,
                   size_t ws_bytes, const int64_t* offsets, int32_t max_count, int32_t* nbr, void* stream) {
    return nl_fill_impl(plan, positions, n_atoms, shifts, n_img, centers, n_centers, cutoff, drop_self, ws, ws_bytes,
                        offsets, max_count, nbr, (hipStream_t)stream);
}

int cusoap_nl_build(const double* positions, int32_t n_atoms, const double* shifts, int32_t n_img,
                    const double* centers, int32_t n_centers, double cutoff, int32_t drop_self, void* tmp, void* ws,
                    size_t ws_bytes, int64_t* offsets, int32_t* nbr, int64_t nbr_capacity, cusoap_nl_plan_t* plan,
                    int64_t* total_host, int32_t* max_count_host, void* stream) {
    const hipStream_t st = (hipStream_t)stream;
    if (!plan || !total_host || !max_count_host) return fail(CUSOAP_ERR_ARG, "nl_build: null argument");
    int rc = nl_plan_impl(centers, n_centers, n_atoms, n_img, cutoff, tmp, plan, st);
    if (rc) return rc;
    if (ws_bytes < plan->ws_bytes) return fail(CUSOAP_ERR_WORKSPACE, "nl_build: workspace smaller than the plan's");
    rc = nl_count_impl(plan, positions, n_atoms, shifts, n_img, centers, n_centers, cutoff, drop_self, ws, ws_bytes,
                       offsets, total_host, max_count_host, st);
    if (rc || *total_host > nbr_capacity) return rc;
    return nl_fill_impl(plan, positions, n_atoms, shifts, n_img, centers, n_centers, cutoff, drop_self, ws, ws_bytes,
                        offsets, *max_count_host, nbr, st);
}

int cusoap_coefficients(const cusoap_basis_t* basis, int32_t precision, const double* positions, int32_t n_atoms,
                        const int32_t* atom_species, const double* shifts, int32_t n_img, const double* centers,
                        int32_t n_centers, const int32_t* center_self_species, const int64_t* offsets,
                        const int32_t* nbr, int32_t create_path, void* coeffs, void* stream) {
    return coefficients_impl(basis, precision, positions, n_atoms, atom_species, shifts, n_img, centers, n_centers,
                             center_self_species, offsets, nbr, create_path, coeffs, (hipStream_t)stream);
}

int cusoap_power_spectrum(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype, int32_t n_centers,
                          const uint32_t* feat_table, int32_t n_feat, void* out, int32_t out_dtype, void* stream) {
    return power_spectrum_impl(basis, coeffs, coeffs_dtype, n_centers, feat_table, n_feat, out, out_dtype,
                               (hipStream_t)stream);
}

int cusoap_cc_projection(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype, int32_t n_centers,
                         void* out, int32_t out_dtype, void* stream) {
    return cc_projection_impl(basis, coeffs, coeffs_dtype, n_centers, out, out_dtype, (hipStream_t)stream);
}

int cusoap_jacobian_workspace_size(const cusoap_basis_t* basis, int32_t mode, int64_t n_edges, int32_t n_centers,
                                   int32_t n_inc, size_t* bytes) {
    if (!basis || !bytes) return fail(CUSOAP_ERR_ARG, "jacobian_workspace_size: null argument");
    *bytes = jacobian_ws_bytes(basis, mode, n_edges, n_centers, n_inc);
    return CUSOAP_OK;
}

int cusoap_jacobian_dense(const cusoap_basis_t* basis, int32_t mode, const double* positions, int32_t n_atoms,
                          const int32_t* atom_species, const double* shifts, int32_t n_img, const double* centers,
                          int32_t n_centers, const int32_t* center_attach, const int64_t* offsets, const int32_t* nbr,
                          int64_t n_edges, const double* coeffs, const int32_t* inv_atom, int32_t n_inc,
                          const uint32_t* feat_table, int32_t n_feat, void* ws, size_t ws_bytes, void* out,
                          int32_t out_dtype, int32_t compute_dtype, void* stream) {
    return jacobian_dense_impl(basis, mode, positions, n_atoms, atom_species, shifts, n_img, centers, n_centers,
                               center_attach, offsets, nbr, n_edges, coeffs, inv_atom, n_inc, feat_table, n_feat, ws,
                               ws_bytes, out, out_dtype, compute_dtype, (hipStream_t)stream);
}

int cusoap_vjp_workspace_size(const cusoap_basis_t* basis, int32_t n_centers, size_t* bytes) {
    if (!basis || !bytes) return fail(CUSOAP_ERR_ARG, "vjp_workspace_size: null argument");
    *bytes = vjp_ws_bytes(basis, n_centers);
    return CUSOAP_OK;
}

int cusoap_vjp_ordered_workspace_size(const cusoap_basis_t* basis, int32_t precision, int32_t n_centers,
                                      int64_t n_edges, size_t* bytes) {
    if (!basis || !bytes) return fail(CUSOAP_ERR_ARG, "vjp_ordered_workspace_size: null argument");
    if (precision < 0 || precision > 2) return fail(CUSOAP_ERR_ARG, "vjp_ordered_workspace_size: unknown precision");
    if (n_centers < 0 || n_edges < 0) return fail(CUSOAP_ERR_ARG, "vjp_ordered_workspace_size: negative size");
    *bytes = vjp_ordered_ws_bytes(basis, precision, n_centers, n_edges);
    return CUSOAP_OK;
}

int cusoap_vjp(const cusoap_basis_t* basis, int32_t mode, int32_t precision, const double* positions,
               int32_t n_atoms, const int32_t* atom_species, const double* shifts, int32_t n_img,
               const double* centers, int32_t n_centers, const int32_t* center_attach,
               const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* coeffs,
               int32_t coeffs_dtype, const int32_t* pair_feature, const void* grad_desc, int32_t gd_dtype,
               int32_t gd_broadcast, int32_t n_feat, void* ws, size_t ws_bytes, double* grad_pos, void* stream) {
    return vjp_impl(basis, mode, precision, positions, n_atoms, atom_species, shifts, n_img, centers, n_centers,
                    center_attach, center_species, offsets, nbr, coeffs, coeffs_dtype, pair_feature, grad_desc,
                    gd_dtype, gd_broadcast, n_feat, ws, ws_bytes, grad_pos, (hipStream_t)stream);
}

int cusoap_edge_reverse_workspace_size(int32_t n_centers, size_t* bytes) {
    if (!bytes || n_centers < 0) return fail(CUSOAP_ERR_ARG, "edge_reverse_workspace_size: invalid argument");
    *bytes = edge_reverse_ws_bytes(n_centers);
    return CUSOAP_OK;
}

int cusoap_edge_reverse(const double* positions, int32_t n_atoms, const double* shifts, int32_t n_img,
                        const double* centers, int32_t n_centers, const int64_t* offsets, const int32_t* nbr,
                        void* ws, size_t ws_bytes, int32_t* rev, void* stream) {
    return edge_reverse_impl(positions, n_atoms, shifts, n_img, centers, n_centers, offsets, nbr, ws, ws_bytes, rev,
                             (hipStream_t)stream);
}

int cusoap_vjp_ordered(const cusoap_basis_t* basis, int32_t mode, int32_t precision, const double* positions,
                       int32_t n_atoms, const int32_t* atom_species, const double* shifts, int32_t n_img,
                       const double* centers, int32_t n_centers, const int32_t* center_attach,
                       const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const int32_t* rev,
                       int64_t n_edges, const void* coeffs, int32_t coeffs_dtype, const int32_t* pair_feature,
                       const void* grad_desc, int32_t gd_dtype, int32_t gd_broadcast, int32_t n_feat, void* ws,
                       size_t ws_bytes, double* grad_pos, void* stream) {
    return vjp_ordered_impl(basis, mode, precision, positions, n_atoms, atom_species, shifts, n_img, centers, n_centers,
                            center_attach, center_species, offsets, nbr, rev, n_edges, coeffs, coeffs_dtype,
                            pair_feature, grad_desc, gd_dtype, gd_broadcast, n_feat, ws, ws_bytes, grad_pos,
                            (hipStream_t)stream);
}

int cusoap_vjp_noalloc(const cusoap_basis_t* basis, int32_t mode, int32_t precision, const double* positions,
                       int32_t n_atoms, const int32_t* atom_species, const double* shifts, int32_t n_img,
                       const double* centers, int32_t n_centers, const int32_t* center_attach,
                       const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* coeffs,
                       int32_t coeffs_dtype, const int32_t* pair_feature, const void* grad_desc, int32_t gd_dtype,
                       int32_t gd_broadcast, int32_t n_feat, double* grad_pos, void* stream) {
    return vjp_noalloc_impl(basis, mode, precision, positions, n_atoms, atom_species, shifts, n_img, centers,
                            n_centers, center_attach, center_species, offsets, nbr, coeffs, coeffs_dtype,
                            pair_feature, grad_desc, gd_dtype, gd_broadcast, n_feat, grad_pos, (hipStream_t)stream);
}

int cusoap_lowdin_invsqrt_f64(const double* S, int32_t batch, int32_t n, double eps, double* out, void* stream) {
    return lowdin_impl(S, batch, n, eps, out, (hipStream_t)stream);
}

int cusoap_power_spectrum_mfma(const cusoap_basis_t* basis, const void* coeffs, int32_t coeffs_dtype,
                               int32_t n_centers, const int32_t* pair_feature, int32_t n_feat, int32_t kind,
                               int32_t split, void* out, int32_t out_dtype, void* stream) {
    return power_spectrum_mfma_impl(basis, coeffs, coeffs_dtype, n_centers, pair_feature, n_feat, kind, split, out,
                                    out_dtype, (hipStream_t)stream);
}

int64_t cusoap_mlp_param_count(int32_t n_hidden, int32_t n_out) { return mlp_param_count(n_hidden, n_out); }

int cusoap_radial_exact(const cusoap_basis_t* basis, const double* r, int32_t n_pts, double* B, double* dBdr,
                        void* stream) {
    return radial_exact_impl(basis, r, n_pts, B, dBdr, (hipStream_t)stream);
}

int cusoap_mlp_eval(const float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out, const float* x,
                    int32_t n_pts, float* y, float* dydx, void* stream) {
    return mlp_eval_impl(params, n_pairs, n_hidden, n_out, x, n_pts, y, dydx, (hipStream_t)stream);
}

int cusoap_mlp_train_workspace_size(int32_t n_pairs, int32_t n_hidden, int32_t n_out, int32_t batch, size_t* bytes) {
    if (!bytes) return fail(CUSOAP_ERR_ARG, "mlp_train_workspace_size: null argument");
    if (mlp_param_count(n_hidden, n_out) < 0) return fail(CUSOAP_ERR_UNSUPPORTED, "mlp: unsupported network shape");
    *bytes = mlp_train_ws_bytes(n_pairs, n_out, batch);
    return CUSOAP_OK;
}

int cusoap_mlp_loss_grad(const float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out, const float* x,
                         const float* t, const float* dt, int32_t n_samples, int64_t target_pair_stride,
                         const float* w_val, const float* w_der, int32_t batch, uint32_t seed, int64_t step, void* ws,
                         size_t ws_bytes, float* grad, float* loss, void* stream) {
    return mlp_loss_grad_impl(params, n_pairs, n_hidden, n_out, x, t, dt, n_samples, target_pair_stride, w_val, w_der,
                              batch, seed, step, ws, ws_bytes, grad, loss, (hipStream_t)stream);
}

int cusoap_adam_step(float* params, const float* grad, float* m, float* v, int64_t n, float lr, float beta1,
                     float beta2, float eps, int64_t step, void* stream) {
    return adam_impl(params, grad, m, v, n, lr, beta1, beta2, eps, step, (hipStream_t)stream);
}

int cusoap_mlp_refit_workspace_size(int32_t n_pairs, int32_t n_hidden, int32_t n_out, int32_t n_samples,
                                    size_t* bytes) {
    if (!bytes) return fail(CUSOAP_ERR_ARG, "mlp_refit_workspace_size: null argument");
    if (mlp_param_count(n_hidden, n_out) < 0) return fail(CUSOAP_ERR_UNSUPPORTED, "mlp: unsupported network shape");
    *bytes = mlp_refit_ws_bytes(n_pairs, n_out, n_samples);
    return CUSOAP_OK;
}

int cusoap_mlp_refit_output(float* params, int32_t n_pairs, int32_t n_hidden, int32_t n_out, const float* x,
                            const float* t, const float* dt, int32_t n_samples, int64_t target_pair_stride,
                            const float* w_der, double ridge, void* ws, size_t ws_bytes, void* stream) {
    return mlp_refit_impl(params, n_pairs, n_hidden, n_out, x, t, dt, n_samples, target_pair_stride, w_der, ridge, ws,
                          ws_bytes, (hipStream_t)stream);
}

int cusoap_radial_table_workspace_size(const cusoap_basis_t* basis, int32_t n_intervals, size_t* bytes) {
    if (!basis || !bytes) return fail(CUSOAP_ERR_ARG, "radial_table_workspace_size: null argument");
    *bytes = radial_table_ws_bytes(basis, n_intervals);
    return CUSOAP_OK;
}

int cusoap_radial_table(const cusoap_basis_t* basis, const float* params, int32_t n_hidden, const int32_t* use_pair,
                        const float* scale, double r_fit, int32_t n_intervals, double r_max, void* ws,
                        size_t ws_bytes, float* table, void* stream) {
    return radial_table_impl(basis, params, n_hidden, use_pair, scale, r_fit, n_intervals, r_max, ws, ws_bytes, table,
                             (hipStream_t)stream);
}

}  // extern "C"
