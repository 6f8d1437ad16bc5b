// Descriptor assembly from the coefficients: per-center power spectrum in the
// DScribe feature order, and the cc projection layout.
//
// Power spectrum (restates _power_spectrum, soap_generator.py:1634-1689):
//   p[c][f] = sqrt(8 pi^2 / (2l+1)) * sum_m c[j][n][lm] c[jd][nd][lm]
// with f -> (j, jd, n, nd, l) from a packed feature table built on the host in
// the order of _build_feature_slices (soap_generator.py:1215-1245).  One
// workgroup per center stages the center's S*n*(l_max+1)^2 coefficients in
// LDS; threads walk the features so the output row is stored coalesced.
//
// cc (restates _projection_cc / _to_dscribe_cc, soap_generator.py:1603-1632):
// l=1 stored as (z, x, y), l>=2 scaled by pi^-3/2.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

__device__ __forceinline__ void unpack_feature(uint32_t v, int& j, int& jd, int& n, int& nd, int& l) {
    j = (int)(v & 0xffu);
    jd = (int)((v >> 8) & 0xffu);
    n = (int)((v >> 16) & 0x3fu);
    nd = (int)((v >> 22) & 0x3fu);
    l = (int)(v >> 28);
}

template <typename TC, typename TO>
__global__ __launch_bounds__(256) void power_spectrum_kernel(int S, int nmax, int lmax,
                                                             const TC* __restrict__ coeffs,
                                                             const uint32_t* __restrict__ ft, int n_feat,
                                                             TO* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int L2 = (lmax + 1) * (lmax + 1);
    const int per = S * nmax * L2;
    TC* cs = (TC*)smem;
    double* pref = (double*)(cs + per + (per & 1));
    const int c = blockIdx.x;
    const TC* src = coeffs + (size_t)c * per;
    for (int i = threadIdx.x; i < per; i += blockDim.x) cs[i] = src[i];
    if ((int)threadIdx.x <= lmax) pref[threadIdx.x] = sqrt(8.0 * M_PI * M_PI / (2 * threadIdx.x + 1));
    __syncthreads();
    TO* dst = out + (size_t)c * n_feat;
    for (int f = threadIdx.x; f < n_feat; f += blockDim.x) {
        int j, jd, n, nd, l;
        unpack_feature(ft[f], j, jd, n, nd, l);
        const TC* a = cs + (j * nmax + n) * L2 + l * l;
        const TC* b = cs + (jd * nmax + nd) * L2 + l * l;
        TC acc = TC(0);
        for (int m = 0; m < 2 * l + 1; ++m) acc += a[m] * b[m];
        dst[f] = (TO)((TC)pref[l] * acc);
    }
}

__device__ __forceinline__ int cc_source(int lmp) {
    // DScribe cc order: output slot lm' reads our slot perm[lm'] (l=1: z, x, y)
    if (lmp == 1) return 2;
    if (lmp == 2) return 3;
    if (lmp == 3) return 1;
    return lmp;
}

template <typename TC, typename TO>
__global__ void cc_kernel(int S, int nmax, int lmax, int64_t n_out, const TC* __restrict__ coeffs,
                          TO* __restrict__ out) {
    const int L2 = (lmax + 1) * (lmax + 1);
    const double s_hi = pow(M_PI, -1.5);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += (int64_t)gridDim.x * blockDim.x) {
        const int lmp = (int)(i % L2);
        const int64_t row = i - lmp;
        const TC v = coeffs[row + cc_source(lmp)];
        out[i] = lmp >= 4 ? (TO)((double)v * s_hi) : (TO)v;
    }
}

template <typename TC, typename TO>
int launch_ps(const cusoap_basis_t& b, const void* coeffs, int n_centers, const uint32_t* ft, int n_feat, void* out,
              hipStream_t st) {
    const int L2 = (b.l_max + 1) * (b.l_max + 1);
    const int per = b.n_species * b.n_max * L2;
    const size_t smem = sizeof(TC) * (size_t)(per + 1) + sizeof(double) * (b.l_max + 2) + 16;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "power_spectrum: too many coefficients per center");
    auto kern = power_spectrum_kernel<TC, TO>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("power_spectrum", st);
    kern<<<n_centers, 256, smem, st>>>(b.n_species, b.n_max, b.l_max, (const TC*)coeffs, ft, n_feat, (TO*)out);
    prof_end("power_spectrum", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <typename TC, typename TO>
int launch_cc(const cusoap_basis_t& b, const void* coeffs, int n_centers, void* out, hipStream_t st) {
    const int L2 = (b.l_max + 1) * (b.l_max + 1);
    const int64_t n_out = (int64_t)n_centers * b.n_species * b.n_max * L2;
    const int64_t blocks = std::min<int64_t>((n_out + 255) / 256, 1 << 20);
    prof_begin("cc_projection", st);
    cc_kernel<TC, TO><<<(unsigned)std::max<int64_t>(blocks, 1), 256, 0, st>>>(b.n_species, b.n_max, b.l_max, n_out,
                                                                              (const TC*)coeffs, (TO*)out);
    prof_end("cc_projection", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace

int power_spectrum_impl(const cusoap_basis_t* b, const void* coeffs, int cdt, int n_centers, const uint32_t* ft,
                        int n_feat, void* out, int odt, hipStream_t st) {
    if (!b || n_feat < 0 || n_centers < 0) return fail(CUSOAP_ERR_ARG, "power_spectrum: invalid argument");
    if (n_centers == 0 || n_feat == 0) return CUSOAP_OK;
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F64) return launch_ps<double, double>(*b, coeffs, n_centers, ft, n_feat, out, st);
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F32) return launch_ps<double, float>(*b, coeffs, n_centers, ft, n_feat, out, st);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F32) return launch_ps<float, float>(*b, coeffs, n_centers, ft, n_feat, out, st);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F64) return launch_ps<float, double>(*b, coeffs, n_centers, ft, n_feat, out, st);
    return fail(CUSOAP_ERR_ARG, "power_spectrum: unknown dtype");
}

int cc_projection_impl(const cusoap_basis_t* b, const void* coeffs, int cdt, int n_centers, void* out, int odt,
                       hipStream_t st) {
    if (!b || n_centers < 0) return fail(CUSOAP_ERR_ARG, "cc_projection: invalid argument");
    if (n_centers == 0) return CUSOAP_OK;
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F64) return launch_cc<double, double>(*b, coeffs, n_centers, out, st);
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F32) return launch_cc<double, float>(*b, coeffs, n_centers, out, st);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F32) return launch_cc<float, float>(*b, coeffs, n_centers, out, st);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F64) return launch_cc<float, double>(*b, coeffs, n_centers, out, st);
    return fail(CUSOAP_ERR_ARG, "cc_projection: unknown dtype");
}

}  // namespace cusoap
