// Adjoint of the descriptor with respect to the positions (vector-Jacobian
// product), without forming d desc / d x.  No reference counterpart: the
// reference only builds the dense Jacobian (soap_generator.py:2270-2626), which
// is infeasible beyond ~10^4 atoms.
//
//   L = sum_{c,f} G_desc[c][f] desc[c][f]
//   dL/dc[c][s][n][lm]:  power spectrum  pref_l sum_{s'n'} W_l[(s n),(s' n')] c[s'][n'][lm]
//                        (W from G_desc through the pair -> feature table, x2 on the diagonal)
//                        cc              G_desc permuted back to (y, z, x) and scaled
//   dL/dx_atom = sum over edges (c -> atom) of sum_{n,lm} dL/dc[c][s][n][lm] g[d][n][lm]
//   with the edge gradient g of the closed-form Jacobian (fixed centers).
//
// Kernel 1 (one workgroup per center) forms dL/dc and, for the GTO basis,
// folds the Loewdin matrix into it: M[s][k][lm] = sum_n beta_l[n][k] dL/dc[s][n][lm],
// so kernel 2 contracts with the primitive radial functions directly.
// Kernel 2 (one wavefront per center, tiles of 64 edges, lane = edge)
// recomputes each edge's solid harmonics (with gradients) and radial factors,
// contracts them with M and scatters the 3-vector with float64 atomics.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

__device__ __forceinline__ int lm_l2(int lm) {
    int l = (int)sqrtf((float)lm);
    while (l * l > lm) --l;
    while ((l + 1) * (l + 1) <= lm) ++l;
    return l;
}

// our lm -> DScribe cc slot holding it (inverse of the (z, x, y) permutation)
__device__ __forceinline__ int cc_slot_of(int lm) {
    if (lm == 1) return 3;
    if (lm == 2) return 1;
    if (lm == 3) return 2;
    return lm;
}

template <typename TG>
__global__ __launch_bounds__(256) void coeff_adjoint_kernel(cusoap_basis_t b, int mode, const double* __restrict__ coeffs,
                                                            const int32_t* __restrict__ pair_feature,
                                                            const TG* __restrict__ gd, int gd_broadcast, int n_feat,
                                                            double* __restrict__ M) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int S = b.n_species, n = b.n_max, L1 = b.l_max + 1, L2 = L1 * L1;
    const int per = S * n * L2, SN = S * n;
    double* cs = (double*)smem;       // [per]
    double* G = cs + per;              // [per]
    double* g = G + per;               // [n_feat]
    const int c = blockIdx.x;
    const double* src = coeffs + (size_t)c * per;
    const TG* gsrc = gd + (gd_broadcast ? 0 : (size_t)c * n_feat);
    for (int i = threadIdx.x; i < per; i += blockDim.x) cs[i] = src[i];
    for (int i = threadIdx.x; i < n_feat; i += blockDim.x) g[i] = (double)gsrc[i];
    __syncthreads();
    for (int t = threadIdx.x; t < per; t += blockDim.x) {
        const int a = t / L2, lm = t - a * L2, l = lm_l2(lm);
        double v = 0.0;
        if (mode == CUSOAP_MODE_PS) {
            const int32_t* pf = pair_feature + ((size_t)l * SN + a) * SN;
            for (int bb = 0; bb < SN; ++bb) {
                const int f = pf[bb];
                if (f >= 0) v += (bb == a ? 2.0 : 1.0) * g[f] * cs[bb * L2 + lm];
            }
            v *= sqrt(8.0 * M_PI * M_PI / (2 * l + 1));
        } else {
            const int slot = cc_slot_of(lm);
            v = g[a * L2 + slot] * (slot >= 4 ? 0.17958712212516656 : 1.0);
        }
        G[t] = v;
    }
    __syncthreads();
    double* Mc = M + (size_t)c * per;
    for (int t = threadIdx.x; t < per; t += blockDim.x) {
        if (b.rbf == CUSOAP_RBF_GTO) {
            const int a = t / L2, lm = t - a * L2, l = lm_l2(lm);
            const int s = a / n, k = a - s * n;
            double v = 0.0;
            for (int nn = 0; nn < n; ++nn) v += b.gto_beta[((size_t)l * n + nn) * n + k] * G[(s * n + nn) * L2 + lm];
            Mc[t] = v;
        } else {
            Mc[t] = G[t];
        }
    }
}

template <int LMAX, int RBF>
__global__ __launch_bounds__(64) void vjp_edge_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                      const int32_t* __restrict__ atom_sp,
                                                      const double* __restrict__ shifts,
                                                      const double* __restrict__ centers,
                                                      const int32_t* __restrict__ center_attach,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ nbr, const double* __restrict__ M,
                                                      double* __restrict__ grad) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const int n = b.n_max, S = b.n_species, NL = n * L1;
    extern __shared__ __align__(16) unsigned char smem[];
    double* Ms = (double*)smem;               // [S][n][L2]
    double* R0 = Ms + S * n * L2;             // [64][NL]
    double* R1 = R0 + 64 * NL;                // [64][NL]
    const int c = blockIdx.x, lane = threadIdx.x;
    const int per = S * n * L2;
    for (int i = lane; i < per; i += 64) Ms[i] = M[(size_t)c * per + i];
    __syncthreads();
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    double* r0 = R0 + lane * NL;
    double* r1 = R1 + lane * NL;
    for (int64_t base = e0; base < e1; base += 64) {
        const int64_t e = base + lane;
        if (e < e1) {
            const int ext = nbr[e];
            const int img = ext / n_atoms, a = ext - img * n_atoms;
            const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
            const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
            const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            const double r = sqrt(r2);
            double w = 1.0, dw = 0.0;
            if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
            const double dwr = r > 0.0 ? dw / r : 0.0;
            const int s = atom_sp[a];
            if (RBF == CUSOAP_RBF_GTO) {
                for (int k = 0; k < n; ++k)
#pragma unroll
                    for (int l = 0; l < L1; ++l) {
                        const double gam = b.gto_gamma[l * n + k];
                        const double q = b.gto_pref[l * n + k] * exp(-gam * r2);
                        r0[k * L1 + l] = w * q;
                        r1[k * L1 + l] = (w * (-2.0 * gam) + dwr) * q;
                    }
            } else {
                for (int k = 0; k < n; ++k) {
                    double B[L1], Bh[L1];
                    if (r < 1e-4) {
#pragma unroll
                        for (int l = 0; l < L1; ++l) {
                            B[l] = b.poly_B0[k * L1 + l];
                            Bh[l] = b.poly_Bh0[k * L1 + l];
                        }
                    } else {
                        double I[L1 + 1], J[L1 + 1];
#pragma unroll
                        for (int l = 0; l <= L1; ++l) I[l] = J[l] = 0.0;
                        const int Q = b.quad_n;
                        poly_levels(b, r, L1, [&](int q, int l, double P) {
                            const double gq = b.poly_gss[k * Q + q];
                            if (l < L1) I[l] += P * gq;
                            if (l >= 1) J[l - 1] += P * gq * b.poly_rx[q];
                        });
                        double rl = 1.0;
#pragma unroll
                        for (int l = 0; l < L1; ++l) {
                            B[l] = I[l] / rl;
                            Bh[l] = (2.0 * b.eta) * (J[l] - r * I[l]) / (rl * r);
                            rl *= r;
                        }
                    }
#pragma unroll
                    for (int l = 0; l < L1; ++l) {
                        r0[k * L1 + l] = w * B[l];
                        r1[k * L1 + l] = w * Bh[l] + dwr * B[l];
                    }
                }
            }
            const double* Mrow = Ms + (size_t)s * n * L2;
            double gx = 0.0, gy = 0.0, gz = 0.0;
            solid_harmonics<LMAX, true, double>(rx, ry, rz, [&](int lm, double y, double yx, double yy, double yz) {
                const int l = lm_l2(lm);
                double h0 = 0.0, h1 = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double mv = Mrow[k * L2 + lm];
                    h0 += r0[k * L1 + l] * mv;
                    h1 += r1[k * L1 + l] * mv;
                }
                const double yh1 = y * h1;
                gx += rx * yh1 + yx * h0;
                gy += ry * yh1 + yy * h0;
                gz += rz * yh1 + yz * h0;
            });
            atomicAdd(&grad[3 * (size_t)a], gx);
            atomicAdd(&grad[3 * (size_t)a + 1], gy);
            atomicAdd(&grad[3 * (size_t)a + 2], gz);
            sx += gx;
            sy += gy;
            sz += gz;
        }
    }
    if (center_attach) {
        const int ca = center_attach[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
            sz += __shfl_xor(sz, o, 64);
        }
        if (ca >= 0 && lane == 0) {
            atomicAdd(&grad[3 * (size_t)ca], -sx);
            atomicAdd(&grad[3 * (size_t)ca + 1], -sy);
            atomicAdd(&grad[3 * (size_t)ca + 2], -sz);
        }
    }
}

template <int LMAX>
int launch_vjp_edges(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                     const double* shifts, const double* centers, int n_centers, const int32_t* center_attach,
                     const int64_t* offsets, const int32_t* nbr, const double* M, double* grad, hipStream_t st) {
    const int L2 = (LMAX + 1) * (LMAX + 1);
    const size_t smem = sizeof(double) * ((size_t)b.n_species * b.n_max * L2 + 2 * 64 * (size_t)b.n_max * (LMAX + 1));
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: LDS tile too large");
    if (b.rbf == CUSOAP_RBF_GTO) {
        auto kern = vjp_edge_kernel<LMAX, CUSOAP_RBF_GTO>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        prof_begin("vjp_edges", st);
        kern<<<n_centers, 64, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, center_attach, offsets, nbr, M, grad);
        prof_end("vjp_edges", st);
    } else {
        auto kern = vjp_edge_kernel<LMAX, CUSOAP_RBF_POLY>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        prof_begin("vjp_edges", st);
        kern<<<n_centers, 64, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, center_attach, offsets, nbr, M, grad);
        prof_end("vjp_edges", st);
    }
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace

size_t vjp_ws_bytes(const cusoap_basis_t* b, int n_centers) {
    const int L2 = (b->l_max + 1) * (b->l_max + 1);
    return align_up(sizeof(double) * (size_t)std::max(n_centers, 1) * b->n_species * b->n_max * L2, 256);
}

int vjp_impl(const cusoap_basis_t* b, int mode, const double* pos, int n_atoms, const int32_t* atom_sp,
             const double* shifts, int n_img, const double* centers, int n_centers, const int32_t* center_attach,
             const int64_t* offsets, const int32_t* nbr, const double* coeffs, const int32_t* pair_feature,
             const void* gd, int gd_dtype, int gd_broadcast, int n_feat, void* ws, size_t ws_bytes, double* grad,
             hipStream_t st) {
    (void)n_img;
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: n_max must be 1..16 and l_max 0..9");
    if (mode == CUSOAP_MODE_PS && !pair_feature) return fail(CUSOAP_ERR_ARG, "vjp: pair_feature table missing");
    if (ws_bytes < vjp_ws_bytes(b, n_centers)) return fail(CUSOAP_ERR_WORKSPACE, "vjp: workspace too small");
    CUSOAP_HIP(hipMemsetAsync(grad, 0, sizeof(double) * 3 * (size_t)n_atoms, st));
    if (n_centers == 0) return CUSOAP_OK;
    const int L2 = (b->l_max + 1) * (b->l_max + 1);
    const int per = b->n_species * b->n_max * L2;
    const size_t smem = sizeof(double) * (2 * (size_t)per + n_feat);
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: descriptor row too large for LDS");
    double* M = (double*)ws;
    if (gd_dtype == CUSOAP_DT_F64) {
        auto kern = coeff_adjoint_kernel<double>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        prof_begin("vjp_adjoint", st);
        kern<<<n_centers, 256, smem, st>>>(*b, mode, coeffs, pair_feature, (const double*)gd, gd_broadcast, n_feat, M);
        prof_end("vjp_adjoint", st);
    } else {
        auto kern = coeff_adjoint_kernel<float>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        prof_begin("vjp_adjoint", st);
        kern<<<n_centers, 256, smem, st>>>(*b, mode, coeffs, pair_feature, (const float*)gd, gd_broadcast, n_feat, M);
        prof_end("vjp_adjoint", st);
    }
    CUSOAP_LAUNCH_CHECK();
#define CUSOAP_CASE(L) \
    case L: return launch_vjp_edges<L>(*b, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach, offsets, nbr, M, grad, st);
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: break;
    }
#undef CUSOAP_CASE
    return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: l_max out of range");
}

// ---------------------------------------------------------------------------
// Loewdin S^-1/2 (float64, cyclic Jacobi; one thread per matrix)
// ---------------------------------------------------------------------------
namespace {
constexpr int kJacobiMaxN = 32;

__global__ void lowdin_kernel(const double* __restrict__ S, int batch, int n, double eps, double* __restrict__ out) {
    const int bidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (bidx >= batch) return;
    double A[kJacobiMaxN * kJacobiMaxN], V[kJacobiMaxN * kJacobiMaxN];
    const double* s = S + (size_t)bidx * n * n;
    for (int i = 0; i < n * n; ++i) A[i] = s[i];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                const double app = A[p * n + p], aqq = A[q * n + q];
                if (fabs(apq) <= 1e-17 * sqrt(fabs(app * aqq)) || apq == 0.0) continue;
                rotated = true;
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
        if (!rotated) break;
    }
    double* o = out + (size_t)bidx * n * n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) {
                const double wk = A[k * n + k] > eps ? A[k * n + k] : eps;
                acc += V[i * n + k] * (1.0 / sqrt(wk)) * V[j * n + k];
            }
            o[i * n + j] = acc;
        }
}
}  // namespace

int lowdin_impl(const double* S, int batch, int n, double eps, double* out, hipStream_t st) {
    if (n < 1 || n > kJacobiMaxN || batch < 0) return fail(CUSOAP_ERR_UNSUPPORTED, "lowdin: n must be 1..32");
    if (batch == 0) return CUSOAP_OK;
    lowdin_kernel<<<(batch + 63) / 64, 64, 0, st>>>(S, batch, n, eps, out);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace cusoap
