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
// Power spectrum on the matrix cores (same features, same order): for each
// center and degree l the Gram matrix G_l = X_l X_l^T of the (S n) x (2l+1)
// coefficient block is formed from 16x16x16 MFMA tiles (fp16 or bf16 operands,
// float32 accumulation), one wavefront per center.  X_l is scaled by a power of
// two so that its largest entry is below 1 (fp16 range); with `split` each
// operand is carried as hi + lo, lo = (x - hi) 2^11 rounded again (fp16; the
// factor keeps the residual out of the subnormal range, bf16 needs none), and
// G = hi hi^T + 2^-11 (hi lo^T + lo hi^T), three products per tile into two
// accumulators.  Entries reach the DScribe feature order through the symmetric
// (l, channel, channel) -> feature table (in LDS), staged in LDS and stored as
// one coalesced row per center.
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

// Vector power spectrum: a grid-stride loop over centers.  The feature table and
// the l prefactors are staged in LDS once per workgroup, and the next center's
// coefficients are loaded into registers while the current center's features
// are formed (thread per feature, m summed in order), so the loop is not one
// global-load latency per center.
constexpr int kPsPre = 8;        // prefetched coefficients per thread (rows up to 2048 values)
template <typename TC, typename TO>
__global__ __launch_bounds__(256) void power_spectrum_kernel(int S, int nmax, int lmax, int n_centers,
                                                             const TC* __restrict__ coeffs,
                                                             const uint32_t* __restrict__ ft, int n_feat,
                                                             TO* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int L2 = (lmax + 1) * (lmax + 1);
    const int per = S * nmax * L2;
    double* pref = (double*)smem;                               // [lmax + 1]
    uint32_t* fts = (uint32_t*)(pref + 16);                     // [n_feat]
    TC* cs = (TC*)(fts + ((n_feat + 3) & ~3));                  // [per]
    const int tid = threadIdx.x;
    for (int f = tid; f < n_feat; f += 256) fts[f] = ft[f];
    if (tid <= lmax) pref[tid] = sqrt(8.0 * M_PI * M_PI / (2 * tid + 1));
    TC pre[kPsPre];
    auto fetch = [&](int c) {
        if (c >= n_centers) return;
        const TC* src = coeffs + (size_t)c * per;
#pragma unroll
        for (int j = 0; j < kPsPre; ++j) {
            const int i = tid + 256 * j;
            if (i < per) pre[j] = src[i];
        }
    };
    fetch(blockIdx.x);
    for (int c = blockIdx.x; c < n_centers; c += gridDim.x) {
#pragma unroll
        for (int j = 0; j < kPsPre; ++j) {
            const int i = tid + 256 * j;
            if (i < per) cs[i] = pre[j];
        }
        const TC* src = coeffs + (size_t)c * per;
        for (int i = 256 * kPsPre + tid; i < per; i += 256) cs[i] = src[i];
        __syncthreads();
        fetch(c + gridDim.x);          // in flight while this center's features are formed
        TO* dst = out + (size_t)c * n_feat;
        for (int f = tid; f < n_feat; f += 256) {
            int j, jd, n, nd, l;
            unpack_feature(fts[f], j, jd, n, nd, l);
            const TC* a = cs + (j * nmax + n) * L2 + l * l;
            const TC* b = cs + (jd * nmax + nd) * L2 + l * l;
            TC acc = TC(0);
            for (int m = 0; m < 2 * l + 1; ++m) acc += a[m] * b[m];
            dst[f] = (TO)((TC)pref[l] * acc);
        }
        __syncthreads();
    }
}

typedef _Float16 mf_h4 __attribute__((ext_vector_type(4)));
typedef short mf_s4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <int KIND>
struct MfmaOp;

template <>
struct MfmaOp<CUSOAP_MFMA_F16> {
    using V = mf_h4;
    static constexpr float kLo = 2048.0f;       // residual pre-scale
    // round to fp16 and return the rounded value as float
    __device__ static float put(V& v, int j, float x) {
        const _Float16 h = (_Float16)x;
        v[j] = h;
        return (float)h;
    }
    __device__ static mf_f4 mma(V a, V b, mf_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
};

template <>
struct MfmaOp<CUSOAP_MFMA_BF16> {
    using V = mf_s4;
    static constexpr float kLo = 1.0f;
    // round to nearest even bf16
    __device__ static float put(V& v, int j, float x) {
        uint32_t u = __float_as_uint(x);
        u += 0x7fffu + ((u >> 16) & 1u);
        const uint32_t hb = u & 0xffff0000u;
        v[j] = (short)(hb >> 16);
        return __uint_as_float(hb);
    }
    __device__ static mf_f4 mma(V a, V b, mf_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
};

// operands of rows [16 t, 16 t + 16) and m in [16 kk, 16 kk + 16) of X_l (scaled):
// lane holds X[16 t + lane % 16][16 kk + 4 (lane / 16) + j], j = 0..3
template <int KIND, bool SPLIT>
__device__ __forceinline__ void mf_operand(const float* cs, int L2, int R, int l, int t, int kk, int lane, float sc,
                                           typename MfmaOp<KIND>::V& hi, typename MfmaOp<KIND>::V& lo) {
    const int a = 16 * t + (lane & 15);
    const int m0 = 16 * kk + 4 * (lane >> 4);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = m0 + j;
        const float x = (a < R && m < 2 * l + 1) ? cs[a * L2 + l * l + m] * sc : 0.0f;
        const float h = MfmaOp<KIND>::put(hi, j, x);
        if (SPLIT) MfmaOp<KIND>::put(lo, j, (x - h) * MfmaOp<KIND>::kLo);
    }
}

constexpr int kPsWaves = 4;

template <typename TC, typename TO, int KIND, bool SPLIT>
__global__ __launch_bounds__(64 * kPsWaves) void ps_mfma_kernel(int S, int nmax, int lmax, int n_centers,
                                                                const TC* __restrict__ coeffs,
                                                                const int32_t* __restrict__ pf, int n_feat,
                                                                TO* __restrict__ out) {
    extern __shared__ __align__(16) float smf[];
    const int L2 = (lmax + 1) * (lmax + 1), R = S * nmax, per = R * L2;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * kPsWaves + wv;
    int* pfs = (int*)smf;                                   // [(lmax+1) R R] feature table
    const int npf = (lmax + 1) * R * R;
    float* cs = smf + npf + (size_t)wv * (per + n_feat);
    float* os = cs + per;
    const bool live = c < n_centers;
    for (int i = threadIdx.x; i < npf; i += blockDim.x) pfs[i] = pf[i];
    if (live) {
        const TC* src = coeffs + (size_t)c * per;
        for (int i = lane; i < per; i += 64) cs[i] = (float)src[i];
        for (int i = lane; i < n_feat; i += 64) os[i] = 0.0f;
    }
    __syncthreads();
    if (live) {
        const int RT = (R + 15) / 16;
        for (int l = 0; l <= lmax; ++l) {
            const int M = 2 * l + 1, KT = (M + 15) / 16;
            float mx = 0.0f;
            for (int i = lane; i < R * M; i += 64) mx = fmaxf(mx, fabsf(cs[(i / M) * L2 + l * l + i % M]));
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            int e = 0;
            if (mx > 0.0f) frexpf(mx, &e);
            const float sc = ldexpf(1.0f, -e);                       // scaled entries lie below 1
            const float back = ldexpf(sqrtf(8.0f * 3.14159265358979f * 3.14159265358979f / (float)M), 2 * e);
            for (int ti = 0; ti < RT; ++ti)
                for (int tj = ti; tj < RT; ++tj) {
                    mf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f}, acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
                    for (int kk = 0; kk < KT; ++kk) {
                        typename MfmaOp<KIND>::V ahi, alo, bhi, blo;
                        mf_operand<KIND, SPLIT>(cs, L2, R, l, ti, kk, lane, sc, ahi, alo);
                        mf_operand<KIND, SPLIT>(cs, L2, R, l, tj, kk, lane, sc, bhi, blo);
                        acc = MfmaOp<KIND>::mma(ahi, bhi, acc);
                        if (SPLIT) {
                            acc2 = MfmaOp<KIND>::mma(ahi, blo, acc2);
                            acc2 = MfmaOp<KIND>::mma(alo, bhi, acc2);
                        }
                    }
                    if (SPLIT) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[j] += acc2[j] * (1.0f / MfmaOp<KIND>::kLo);
                    }
                    // lane holds G[16 ti + 4 (lane / 16) + j][16 tj + lane % 16]
                    const int b = 16 * tj + (lane & 15);
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int a = 16 * ti + 4 * (lane >> 4) + j;
                        if (a < R && b < R && (ti < tj || a <= b)) {
                            const int f = pfs[(l * R + a) * R + b];
                            if (f >= 0) os[f] = back * acc[j];
                        }
                    }
                }
        }
    }
    __syncthreads();
    if (live) {
        TO* dst = out + (size_t)c * n_feat;
        for (int i = lane; i < n_feat; i += 64) dst[i] = (TO)os[i];
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
    const size_t smem = sizeof(double) * 16 + sizeof(uint32_t) * (size_t)((n_feat + 3) & ~3) + sizeof(TC) * (size_t)per;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "power_spectrum: too many coefficients per center");
    auto kern = power_spectrum_kernel<TC, TO>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    const int grid = n_centers < 2048 ? n_centers : 2048;
    prof_begin("power_spectrum", st);
    kern<<<grid, 256, smem, st>>>(b.n_species, b.n_max, b.l_max, n_centers, (const TC*)coeffs, ft, n_feat, (TO*)out);
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

template <typename TC, typename TO>
int launch_ps_mfma(const cusoap_basis_t& b, const void* coeffs, int n_centers, const int32_t* pf, int n_feat,
                   int kind, int split, void* out, hipStream_t st) {
    const int L2 = (b.l_max + 1) * (b.l_max + 1);
    const int per = b.n_species * b.n_max * L2;
    const int R = b.n_species * b.n_max;
    const size_t smem = sizeof(float) * ((size_t)kPsWaves * (per + n_feat) + (size_t)(b.l_max + 1) * R * R);
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "power_spectrum_mfma: descriptor row too large for LDS");
    const unsigned blocks = (unsigned)((n_centers + kPsWaves - 1) / kPsWaves);
#define CUSOAP_PSM(K, SP)                                                                                      \
    do {                                                                                                       \
        auto kern = ps_mfma_kernel<TC, TO, K, SP>;                                                             \
        if (smem > 64 * 1024)                                                                                  \
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem)); \
        kern<<<blocks, 64 * kPsWaves, smem, st>>>(b.n_species, b.n_max, b.l_max, n_centers, (const TC*)coeffs, pf,  \
                                                   n_feat, (TO*)out);                                         \
    } while (0)
    prof_begin("power_spectrum", st);
    if (kind == CUSOAP_MFMA_F16) {
        if (split) CUSOAP_PSM(CUSOAP_MFMA_F16, true);
        else CUSOAP_PSM(CUSOAP_MFMA_F16, false);
    } else {
        if (split) CUSOAP_PSM(CUSOAP_MFMA_BF16, true);
        else CUSOAP_PSM(CUSOAP_MFMA_BF16, false);
    }
#undef CUSOAP_PSM
    prof_end("power_spectrum", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace

int power_spectrum_mfma_impl(const cusoap_basis_t* b, const void* coeffs, int cdt, int n_centers, const int32_t* pf,
                             int n_feat, int kind, int split, void* out, int odt, hipStream_t st) {
    if (!b || !pf || n_feat < 0 || n_centers < 0) return fail(CUSOAP_ERR_ARG, "power_spectrum_mfma: invalid argument");
    if (kind != CUSOAP_MFMA_F16 && kind != CUSOAP_MFMA_BF16)
        return fail(CUSOAP_ERR_ARG, "power_spectrum_mfma: kind must be CUSOAP_MFMA_F16 or CUSOAP_MFMA_BF16");
    if (n_centers == 0 || n_feat == 0) return CUSOAP_OK;
#define CUSOAP_PSA(TC, TO) launch_ps_mfma<TC, TO>(*b, coeffs, n_centers, pf, n_feat, kind, split, out, st)
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F64) return CUSOAP_PSA(double, double);
    if (cdt == CUSOAP_DT_F64 && odt == CUSOAP_DT_F32) return CUSOAP_PSA(double, float);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F32) return CUSOAP_PSA(float, float);
    if (cdt == CUSOAP_DT_F32 && odt == CUSOAP_DT_F64) return CUSOAP_PSA(float, double);
#undef CUSOAP_PSA
    return fail(CUSOAP_ERR_ARG, "power_spectrum_mfma: unknown dtype");
}

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
