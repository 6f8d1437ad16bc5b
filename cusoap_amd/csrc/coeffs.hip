// Fused density-coefficient kernel c[center][s][n][lm].
//
// Restates _coefficients_gto / _coefficients_polynomial / _add_self_terms
// (soap_generator.py:1294-1435, 821-850) and the Cacc accumulation of the
// closed-form derivative path (soap_generator.py:2414-2423) as one pass per
// center:
//
//   one 256-thread workgroup per center walks the center's CSR edge segment
//   in tiles of 64 edges.  Phase A evaluates the tile: thread per edge for the
//   real solid harmonics r^l Y_lm, thread per (edge, radial item) for the
//   radial factors, both stored in LDS as [value][edge] rows (padded to 65 so
//   that threads reading different rows hit different banks).  Phase B:
//   wavefront w takes edges w, w+4, ... of the tile; each lane owns register
//   accumulators for (2 radial n) x (4 consecutive lm of one l) x (up to 4
//   species, chosen by a wavefront-uniform branch on the edge's species).
//   The four wavefronts' partial sums are combined in LDS in a fixed order, so
//   results are bitwise reproducible.
//
// Arithmetic by precision:
//   F64    harmonics, radial, accumulation in double; GTO accumulates the
//          primitives w pref_kl exp(-gamma_kl r^2) and applies the Loewdin
//          beta_l once per center (S n^2 L^2 work instead of per edge).
//   F32    the same in float (the reference's algorithm in single precision:
//          the beta contraction then loses digits at large n_max).
//   MIXED  radial functions in double with beta applied per edge, so the
//          orthonormal B_nl(r) (O(1), no cancellation left) is rounded to
//          float once; harmonics and accumulation in float.
// Polynomial basis: the quadrature radial integral is evaluated in double and
// already orthonormal (no beta).
// Radial table (basis.rad_table, the learned radial model): B_nl(r) by cubic
// Hermite interpolation of the tabulated values and slopes, already orthonormal;
// float accumulation except for F64.
//
// create_path: edges with |r| <= 1e-12 are skipped and the l=0 self term is
// added for index centers, as the reference does; otherwise every edge counts.
#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kRbfTable = 2;     // kernel variant: tabulated radial model
constexpr int kTE = 64;          // edges per tile
constexpr int kTEP = kTE + 1;    // padded LDS row length
constexpr int kThreads = 256;

__host__ __device__ constexpr int n_groups(int lmax) {
    int g = 0;
    for (int l = 0; l <= lmax; ++l) g += (2 * l + 1 + 3) / 4;
    return g;
}

template <int PREC>
struct CT;
template <>
struct CT<CUSOAP_PREC_F64> {
    using TY = double;
    using TA = double;
    using TQ = double;
};
template <>
struct CT<CUSOAP_PREC_F32> {
    using TY = float;
    using TA = float;
    using TQ = float;
};
template <>
struct CT<CUSOAP_PREC_MIXED> {
    using TY = float;
    using TA = float;
    using TQ = double;
};

template <typename T>
__device__ __forceinline__ T t_exp(T v);
template <>
__device__ __forceinline__ float t_exp<float>(float v) { return expf(v); }
template <>
__device__ __forceinline__ double t_exp<double>(double v) { return exp(v); }

template <int LMAX, int PREC, int RBF>
__global__ __launch_bounds__(kThreads) void coeff_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                         const int32_t* __restrict__ atom_sp,
                                                         const double* __restrict__ shifts,
                                                         const double* __restrict__ centers,
                                                         const int32_t* __restrict__ self_sp,
                                                         const int64_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ nbr, int create_path,
                                                         typename CT<PREC>::TA* __restrict__ out) {
    using TY = typename CT<PREC>::TY;
    using TA = typename CT<PREC>::TA;
    using TQ = typename CT<PREC>::TQ;
    constexpr bool kBetaPerCenter = (RBF == CUSOAP_RBF_GTO) && (PREC != CUSOAP_PREC_MIXED);
    constexpr int L1 = LMAX + 1, L2 = L1 * L1, NG = n_groups(LMAX);
    const int n = b.n_max, S = b.n_species, NL = n * L1;
    extern __shared__ __align__(16) unsigned char smem[];
    double* geo = (double*)smem;                   // [kTE][8]: rx ry rz r2 r w
    int* spe = (int*)(geo + kTE * 8);              // [kTE] species, -1 = skip
    TA* Rs = (TA*)(spe + kTE);                     // [NL][kTEP]
    TY* Ys = (TY*)(Rs + NL * kTEP + 2);            // [L2][kTEP]

    const int c = blockIdx.x, tid = threadIdx.x;
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int wv = tid >> 6, lane = tid & 63;
    const int npair = (n + 1) / 2;
    const int nslot_all = npair * NG;                 // (n pair, lm group) slots per species
    const int self_s = (create_path && self_sp) ? self_sp[c] : -1;
    const int cslot = (RBF == kRbfTable && self_sp[c] >= 0) ? self_sp[c] : S;   // table slot of the center

    for (int s0 = 0; s0 < S; s0 += 4) {
        const int s1 = (s0 + 4 < S) ? s0 + 4 : S;
        for (int q0 = 0; q0 < nslot_all; q0 += 64) {
            // ---- this lane's slot: n pair p (n0 = 2p, n1 = 2p+1) x lm group g of degree l
            const int q = q0 + lane;
            const bool valid = q < nslot_all;
            int n0 = 0, l = 0, lm0 = 0, cnt = 0;
            bool has1 = false;
            if (valid) {
                const int pp = q / NG;
                int g = q - pp * NG;
                while (g >= (2 * l + 1 + 3) / 4) {
                    g -= (2 * l + 1 + 3) / 4;
                    ++l;
                }
                n0 = 2 * pp;
                has1 = n0 + 1 < n;
                lm0 = l * l + 4 * g;
                const int left = 2 * l + 1 - 4 * g;
                cnt = left < 4 ? left : 4;
            }
            const int r0i = (n0 * L1 + l) * kTEP, r1i = ((n0 + 1) * L1 + l) * kTEP, yi = lm0 * kTEP;
            TA acc[4][2][4];
#pragma unroll
            for (int a4 = 0; a4 < 4; ++a4)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[a4][h][j] = TA(0);

            for (int64_t t0 = e0; t0 < e1; t0 += kTE) {
                const int nt = (int)((e1 - t0) < kTE ? (e1 - t0) : kTE);
                // ---- A1: geometry
                if (tid < nt) {
                    const int ext = nbr[t0 + tid];
                    const int img = ext / n_atoms, a = ext - img * n_atoms;
                    const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
                    const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
                    const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
                    const double r2 = (rx * rx + ry * ry) + rz * rz;
                    const double r = sqrt(r2);
                    double w = 1.0, dw = 0.0;
                    if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
                    double* g = geo + tid * 8;
                    g[0] = rx; g[1] = ry; g[2] = rz; g[3] = r2; g[4] = r; g[5] = w;
                    const int sp = atom_sp[a];
                    spe[tid] = ((create_path && !(r > 1e-12)) || sp < s0 || sp >= s1) ? -1 : sp;
                }
                __syncthreads();
                // ---- A2: harmonics (item per edge) and radial factors
                int n_rad;
                if (RBF == CUSOAP_RBF_POLY) n_rad = nt * n;
                else if (RBF == kRbfTable) n_rad = nt * L1;
                else if (kBetaPerCenter) n_rad = nt * NL;
                else n_rad = nt * L1;
                for (int item = tid; item < nt + n_rad; item += kThreads) {
                    if (item < nt) {
                        if (spe[item] < 0) continue;
                        const double* g = geo + item * 8;
                        TY* y = Ys + item;
                        solid_harmonics<LMAX, false, TY>((TY)g[0], (TY)g[1], (TY)g[2],
                                                         [&](int lm, TY v, TY, TY, TY) { y[lm * kTEP] = v; });
                        continue;
                    }
                    const int it = item - nt;
                    if (RBF == kRbfTable) {
                        const int ll = it / nt, e = it - ll * nt;
                        if (spe[e] < 0) continue;
                        const int G = b.rad_intervals;
                        const double xr = geo[e * 8 + 4] * ((double)G / b.rad_rmax);
                        int i = (int)xr;
                        i = i < G - 1 ? i : G - 1;
                        const TA t = (TA)(xr - i), hh = (TA)(b.rad_rmax / G);
                        const TA t2 = t * t, t3 = t2 * t;
                        const TA h00 = TA(2) * t3 - TA(3) * t2 + TA(1), h01 = TA(3) * t2 - TA(2) * t3;
                        const TA h10 = hh * (t3 - TA(2) * t2 + t), h11 = hh * (t3 - t2);
                        const int NS = L1 * 2 * n;
                        const float* p0 = b.rad_table + ((size_t)(cslot * S + spe[e]) * (G + 1) + i) * NS + ll * 2 * n;
                        const float* p1 = p0 + NS;
                        const TA w = (TA)geo[e * 8 + 5];
                        for (int nn = 0; nn < n; ++nn) {
                            const TA B = h00 * (TA)p0[nn] + h10 * (TA)p0[n + nn] + h01 * (TA)p1[nn] + h11 * (TA)p1[n + nn];
                            Rs[(nn * L1 + ll) * kTEP + e] = w * B;
                        }
                    } else if (RBF == CUSOAP_RBF_GTO && kBetaPerCenter) {
                        const int t = it / nt, e = it - t * nt;
                        if (spe[e] < 0) continue;
                        const int k = t / L1, ll = t - k * L1;
                        const TQ pref = (TQ)b.gto_pref[ll * n + k], gam = (TQ)b.gto_gamma[ll * n + k];
                        Rs[t * kTEP + e] = (TA)((TQ)geo[e * 8 + 5] * (pref * t_exp<TQ>(-gam * (TQ)geo[e * 8 + 3])));
                    } else if (RBF == CUSOAP_RBF_GTO) {
                        const int ll = it / nt, e = it - ll * nt;
                        if (spe[e] < 0) continue;
                        const double r2 = geo[e * 8 + 3], w = geo[e * 8 + 5];
                        double qv[kMaxN];
#pragma unroll
                        for (int k = 0; k < kMaxN; ++k)
                            qv[k] = k < n ? b.gto_pref[ll * n + k] * exp(-b.gto_gamma[ll * n + k] * r2) : 0.0;
                        for (int nn = 0; nn < n; ++nn) {
                            const double* beta = b.gto_beta + ((size_t)ll * n + nn) * n;
                            double B = 0.0;
#pragma unroll
                            for (int k = 0; k < kMaxN; ++k)
                                if (k < n) B += beta[k] * qv[k];
                            Rs[(nn * L1 + ll) * kTEP + e] = (TA)(w * B);
                        }
                    } else {
                        const int k = it / nt, e = it - k * nt;
                        if (spe[e] < 0) continue;
                        const double r = geo[e * 8 + 4], w = geo[e * 8 + 5];
                        double I[L1];
#pragma unroll
                        for (int ll = 0; ll < L1; ++ll) I[ll] = 0.0;
                        if (!create_path && r < 1e-4) {
#pragma unroll
                            for (int ll = 0; ll < L1; ++ll) I[ll] = b.poly_B0[k * L1 + ll];
                        } else {
                            const int Q = b.quad_n;
                            poly_levels(b, r, LMAX, [&](int qq, int ll, double P) { I[ll] += P * b.poly_gss[k * Q + qq]; });
                            double rl = 1.0;
#pragma unroll
                            for (int ll = 0; ll < L1; ++ll) {
                                I[ll] /= rl;
                                rl *= r;
                            }
                        }
#pragma unroll
                        for (int ll = 0; ll < L1; ++ll) Rs[(k * L1 + ll) * kTEP + e] = (TA)(w * I[ll]);
                    }
                }
                __syncthreads();
                // ---- B: wavefront wv sweeps edges wv, wv+4, ...; species is wavefront-uniform
                if (valid) {
                    for (int e = wv; e < nt; e += 4) {
                        const int se = __builtin_amdgcn_readfirstlane(spe[e]);
                        if (se < 0) continue;
                        const TA rv0 = Rs[r0i + e];
                        const TA rv1 = has1 ? Rs[r1i + e] : TA(0);
                        TA yv[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) yv[j] = j < cnt ? (TA)Ys[yi + j * kTEP + e] : TA(0);
                        const int sl = se - s0;
#pragma unroll
                        for (int a4 = 0; a4 < 4; ++a4) {
                            if (sl == a4) {
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    acc[a4][0][j] += rv0 * yv[j];
                                    acc[a4][1][j] += rv1 * yv[j];
                                }
                            }
                        }
                    }
                }
                __syncthreads();
            }
            // ---- combine the four wavefronts' partial sums in a fixed order (LDS, aliases the tile)
            TA* part = Rs;                                  // [4 species][n][L2]
            for (int w2 = 0; w2 < 4; ++w2) {
                if (wv == w2 && valid) {
#pragma unroll
                    for (int a4 = 0; a4 < 4; ++a4) {
                        if (s0 + a4 >= s1) continue;
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            if (h == 1 && !has1) continue;
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                if (j >= cnt) continue;
                                TA* dst = part + (a4 * n + n0 + h) * L2 + lm0 + j;
                                *dst = (w2 == 0) ? acc[a4][h][j] : *dst + acc[a4][h][j];
                            }
                        }
                    }
                }
                __syncthreads();
            }
            // ---- write this pass's outputs (raw primitive sums when beta is applied per center)
            if (wv == 0 && valid) {
                for (int a4 = 0; a4 < s1 - s0; ++a4)
                    for (int h = 0; h < 2; ++h) {
                        if (h == 1 && !has1) continue;
                        const int nn = n0 + h;
                        TA* dst = out + (((size_t)c * S + s0 + a4) * n + nn) * L2 + lm0;
                        for (int j = 0; j < cnt; ++j) {
                            TA v = part[(a4 * n + nn) * L2 + lm0 + j];
                            if (!kBetaPerCenter && s0 + a4 == self_s && lm0 + j == 0) v += (TA)b.self_l0[nn];
                            dst[j] = v;
                        }
                    }
            }
            __syncthreads();
        }
    }
    if (kBetaPerCenter) {
        // ---- Loewdin contraction of the primitive sums, once per center
        __threadfence_block();
        TA* Cs = Rs;                                        // [S][n][L2]
        const int per = S * n * L2;
        TA* oc = out + (size_t)c * per;
        for (int i = tid; i < per; i += kThreads) Cs[i] = oc[i];
        __syncthreads();
        for (int o = tid; o < per; o += kThreads) {
            const int s = o / (n * L2), rem = o - s * n * L2;
            const int nn = rem / L2, lm = rem - nn * L2;
            int ll = 0;
            while ((ll + 1) * (ll + 1) <= lm) ++ll;
            const double* beta = b.gto_beta + ((size_t)ll * n + nn) * n;
            TA v = TA(0);
            for (int k = 0; k < n; ++k) v += (TA)beta[k] * Cs[(s * n + k) * L2 + lm];
            if (s == self_s && lm == 0) v += (TA)b.self_l0[nn];
            oc[o] = v;
        }
    }
}

template <int LMAX, int PREC, int RBF>
int launch_coeff(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                 const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                 const int64_t* offsets, const int32_t* nbr, int create_path, void* out, hipStream_t st) {
    using TY = typename CT<PREC>::TY;
    using TA = typename CT<PREC>::TA;
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const int n = b.n_max, NL = n * L1;
    const size_t tile = sizeof(double) * kTE * 8 + sizeof(int) * kTE + sizeof(TA) * ((size_t)NL * kTEP + 2) +
                        sizeof(TY) * (size_t)L2 * kTEP;
    const size_t cs = sizeof(TA) * (size_t)std::max(b.n_species, 4) * n * L2 + sizeof(double) * kTE * 8 +
                      sizeof(int) * kTE;
    const size_t smem = tile > cs ? tile : cs;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: LDS tile too large");
    auto kern = coeff_kernel<LMAX, PREC, RBF>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("coefficients", st);
    kern<<<n_centers, kThreads, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, self_sp, offsets, nbr,
                                            create_path, (TA*)out);
    prof_end("coefficients", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <int LMAX>
int dispatch_prec(const cusoap_basis_t& b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                  const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                  const int64_t* offsets, const int32_t* nbr, int create_path, void* out, hipStream_t st) {
#define CUSOAP_COEFF_ARGS b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st
    if (b.rad_table) {
        if (prec == CUSOAP_PREC_F64) return launch_coeff<LMAX, CUSOAP_PREC_F64, kRbfTable>(CUSOAP_COEFF_ARGS);
        return launch_coeff<LMAX, CUSOAP_PREC_MIXED, kRbfTable>(CUSOAP_COEFF_ARGS);
    }
    if (b.rbf == CUSOAP_RBF_GTO) {
        if (prec == CUSOAP_PREC_F64) return launch_coeff<LMAX, CUSOAP_PREC_F64, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
        if (prec == CUSOAP_PREC_F32) return launch_coeff<LMAX, CUSOAP_PREC_F32, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
        return launch_coeff<LMAX, CUSOAP_PREC_MIXED, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
    }
    if (prec == CUSOAP_PREC_F64) return launch_coeff<LMAX, CUSOAP_PREC_F64, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
    if (prec == CUSOAP_PREC_F32) return launch_coeff<LMAX, CUSOAP_PREC_F32, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
    return launch_coeff<LMAX, CUSOAP_PREC_MIXED, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
#undef CUSOAP_COEFF_ARGS
}

}  // namespace

int coefficients_impl(const cusoap_basis_t* b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                      const double* shifts, int n_img, const double* centers, int n_centers,
                      const int32_t* self_sp, const int64_t* offsets, const int32_t* nbr, int create_path,
                      void* out, hipStream_t st) {
    (void)n_img;
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL || b->n_species < 1)
        return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: n_max must be 1..16 and l_max 0..9");
    if (prec < 0 || prec > 2) return fail(CUSOAP_ERR_ARG, "coefficients: unknown precision");
    if (create_path && !self_sp) return fail(CUSOAP_ERR_ARG, "coefficients: create path needs center_self_species");
    if (b->rad_table && (!self_sp || b->rad_intervals < 1 || !(b->rad_rmax > 0.0)))
        return fail(CUSOAP_ERR_ARG, "coefficients: radial table needs center_self_species and a valid grid");
    if (n_centers == 0) return CUSOAP_OK;
#define CUSOAP_CASE(L) \
    case L: return dispatch_prec<L>(*b, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st);
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: break;
    }
#undef CUSOAP_CASE
    return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: l_max out of range");
}

}  // namespace cusoap
