// Dense analytical Jacobian d desc[c][f] / d x[a][d], layout (C, n_inc, 3, n_feat).
//
// Restates derivatives_analytical_ps (soap_generator.py:2270-2626) -- fixed
// centers, self edges kept -- and generalises it to what the reference gets
// from autograd in derivatives_analytical (soap_generator.py:2007-2113):
// several periodic images of one atom add into one row, the weighting slope
// enters the edge gradient, and with attach an index center's own row also
// receives minus the sum of all its edge gradients.
//
// Edge gradient of c[s][n][lm] (edge species s):
//   g[d][n][lm] = r_d * Bt_nl * Y_lm + Bw_nl * dY_lm/dr_d,
//   Bw = w B,  Bt = w Bhat + (w'/r) B,   Bhat = (1/r) dB/dr.
// Power spectrum, per (center, row) and feature (j,n | jd,nd | l):
//   pref_l * sum_edges [ d_{j,s}(r_d Bt_n U[jd nd] + Bw_n V_d[jd nd])
//                      + d_{jd,s}(r_d Bt_nd U[j n] + Bw_nd V_d[j n]) ]
// with per-edge contractions U[kq][l] = sum_m Y_lm c[kq][lm] and
// V_d[kq][l] = sum_m dY_lm/dr_d c[kq][lm].
//
// Kernel 1 (edge_kernel) writes one float64 record per edge:
//   rvec[3] | Bw[n][l] | Bt[n][l] | ps: U[S n][l], V[3][S n][l]  (cc: Y[L2], dY[3][L2])
// Kernel 2 (dense_kernel) takes one center per workgroup, sorts the center's
// included edges by output row in LDS, and walks the rows in order: runs of
// rows with no edge are stored as flat 16-byte zero writes, rows with edges
// are computed from the records.  Every output element is stored exactly once.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kJacCap = 4096;   // max included edges per center in the dense kernel

struct RecLayout {
    int NL, SNL, L1, L2, stride, offB, offBt, offU, offV;
};

__host__ __device__ inline RecLayout rec_layout(int S, int n, int lmax, int mode) {
    RecLayout r;
    r.L1 = lmax + 1;
    r.L2 = r.L1 * r.L1;
    r.NL = n * r.L1;
    r.SNL = S * r.NL;
    r.offB = 3;
    r.offBt = 3 + r.NL;
    r.offU = 3 + 2 * r.NL;
    if (mode == CUSOAP_MODE_PS) {
        r.offV = r.offU + r.SNL;
        r.stride = r.offV + 3 * r.SNL;
    } else {
        r.offV = r.offU + r.L2;
        r.stride = r.offV + 3 * r.L2;
    }
    r.stride += r.stride & 1;
    return r;
}

__device__ __forceinline__ int lm_l(int lm) {
    int l = (int)sqrtf((float)lm);
    while (l * l > lm) --l;
    while ((l + 1) * (l + 1) <= lm) ++l;
    return l;
}

__device__ __forceinline__ int cc_src(int lmp) {
    if (lmp == 1) return 2;
    if (lmp == 2) return 3;
    if (lmp == 3) return 1;
    return lmp;
}

// ---------------------------------------------------------------------------
// Kernel 1: per-edge records.  4 wavefronts per workgroup, one edge per wave.
// ---------------------------------------------------------------------------
template <int LMAX, int RBF>
__global__ __launch_bounds__(256) void edge_kernel(cusoap_basis_t b, int mode, const double* __restrict__ pos,
                                                   int n_atoms, const double* __restrict__ shifts,
                                                   const double* __restrict__ centers,
                                                   const int64_t* __restrict__ offsets,
                                                   const int32_t* __restrict__ nbr,
                                                   const double* __restrict__ coeffs, double* __restrict__ ws) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const int n = b.n_max, S = b.n_species;
    const RecLayout R = rec_layout(S, n, LMAX, mode);
    const int NL = R.NL;
    extern __shared__ __align__(16) unsigned char smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per_wave = 4 * L2 + 4 * NL;
    double* Ysh = (double*)smem + wv * per_wave;   // [4][L2]: Y, dY/dx, dY/dy, dY/dz
    double* Qs = Ysh + 4 * L2;                     // [NL] primitives (GTO) / B (poly)
    double* Qh = Qs + NL;                          // [NL]
    double* Bw = Qh + NL;                          // [NL]
    double* Bt = Bw + NL;                          // [NL]

    const int c = blockIdx.x;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int64_t ne = e1 - e0;
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const double* cc = coeffs + (size_t)c * S * n * L2;
    const int64_t iters = (ne + 3) / 4;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t e = e0 + it * 4 + wv;
        const bool active = e < e1;
        double rx = 0, ry = 0, rz = 0, r2 = 0, r = 0, w = 1.0, dw = 0.0;
        double* rec = ws + (size_t)(active ? e : 0) * R.stride;
        if (active) {
            const int ext = nbr[e];
            const int img = ext / n_atoms, a = ext - img * n_atoms;
            rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
            ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
            rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
            r2 = (rx * rx + ry * ry) + rz * rz;
            r = sqrt(r2);
            if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
            if (lane == 0) {
                rec[0] = rx;
                rec[1] = ry;
                rec[2] = rz;
                solid_harmonics<LMAX, true, double>(rx, ry, rz, [&](int lm, double v, double gx, double gy, double gz) {
                    Ysh[lm] = v;
                    Ysh[L2 + lm] = gx;
                    Ysh[2 * L2 + lm] = gy;
                    Ysh[3 * L2 + lm] = gz;
                });
            }
            if (RBF == CUSOAP_RBF_GTO) {
                for (int t = lane; t < NL; t += 64) {
                    const int k = t / L1, l = t - k * L1;
                    const double g = b.gto_gamma[l * n + k];
                    const double q = b.gto_pref[l * n + k] * exp(-g * r2);
                    Qs[t] = q;
                    Qh[t] = -2.0 * g * q;
                }
            } else {
                for (int k = lane; k < n; k += 64) {
                    double I[L1 + 1], J[L1 + 1];
                    if (r < 1e-4) {
#pragma unroll
                        for (int l = 0; l < L1; ++l) {
                            Qs[k * L1 + l] = b.poly_B0[k * L1 + l];
                            Qh[k * L1 + l] = b.poly_Bh0[k * L1 + l];
                        }
                    } else {
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
                            Qs[k * L1 + l] = I[l] / rl;
                            Qh[k * L1 + l] = (2.0 * b.eta) * (J[l] - r * I[l]) / (rl * r);
                            rl *= r;
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (active) {
            const double dwr = r > 0.0 ? dw / r : 0.0;
            for (int t = lane; t < NL; t += 64) {
                double B, Bh;
                if (RBF == CUSOAP_RBF_GTO) {
                    const int nn = t / L1, l = t - nn * L1;
                    const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
                    B = 0.0;
                    Bh = 0.0;
                    for (int k = 0; k < n; ++k) {
                        B += beta[k] * Qs[k * L1 + l];
                        Bh += beta[k] * Qh[k * L1 + l];
                    }
                } else {
                    B = Qs[t];
                    Bh = Qh[t];
                }
                const double bw = w * B, bt = w * Bh + dwr * B;
                Bw[t] = bw;
                Bt[t] = bt;
                rec[R.offB + t] = bw;
                rec[R.offBt + t] = bt;
            }
        }
        __syncthreads();
        if (active) {
            if (mode == CUSOAP_MODE_PS) {
                const int SNL = R.SNL;
                for (int t = lane; t < 4 * SNL; t += 64) {
                    const int comp = t / SNL, rem = t - comp * SNL;
                    const int kq = rem / L1, l = rem - kq * L1;
                    const double* y = Ysh + comp * L2 + l * l;
                    const double* cv = cc + (size_t)kq * L2 + l * l;
                    double acc = 0.0;
                    for (int m = 0; m < 2 * l + 1; ++m) acc += y[m] * cv[m];
                    rec[R.offU + t] = acc;   // U then V_x, V_y, V_z (offV = offU + SNL)
                }
            } else {
                for (int t = lane; t < 4 * L2; t += 64) rec[R.offU + t] = Ysh[t];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Kernel 2: dense rows
// ---------------------------------------------------------------------------
template <typename TO>
__device__ __forceinline__ void zero_fill(TO* p, size_t cnt) {
    if (cnt == 0) return;
    unsigned char* bp = (unsigned char*)p;
    const size_t bytes = cnt * sizeof(TO);
    size_t head = (16 - ((uintptr_t)bp & 15)) & 15;
    if (head > bytes) head = bytes;
    const size_t nhead = head / sizeof(TO);
    for (size_t i = threadIdx.x; i < nhead; i += blockDim.x) p[i] = TO(0);
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4* mid = (u32x4*)(bp + head);
    const size_t nmid = (bytes - head) / 16;
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (size_t i = threadIdx.x; i < nmid; i += blockDim.x) __builtin_nontemporal_store(z, &mid[i]);
    TO* tp = (TO*)(bp + head + nmid * 16);
    const size_t ntail = (bytes - head - nmid * 16) / sizeof(TO);
    for (size_t i = threadIdx.x; i < ntail; i += blockDim.x) tp[i] = TO(0);
}

struct FeatDec {
    int j, jd, n, nd, l;
};

__device__ __forceinline__ FeatDec dec(uint32_t v) {
    FeatDec f;
    f.j = (int)(v & 0xffu);
    f.jd = (int)((v >> 8) & 0xffu);
    f.n = (int)((v >> 16) & 0x3fu);
    f.nd = (int)((v >> 22) & 0x3fu);
    f.l = (int)(v >> 28);
    return f;
}

__device__ __forceinline__ double ps_term(const double* rec, const RecLayout& R, int nmax, int s, int d,
                                          const FeatDec& f) {
    const double rv = rec[d];
    const double* B = rec + R.offB;
    const double* Bt = rec + R.offBt;
    const double* U = rec + R.offU;
    const double* V = rec + R.offV + (size_t)d * R.SNL;
    double v = 0.0;
    if (f.j == s) {
        const int a = f.n * R.L1 + f.l, q = (f.jd * nmax + f.nd) * R.L1 + f.l;
        v += rv * Bt[a] * U[q] + B[a] * V[q];
    }
    if (f.jd == s) {
        const int a = f.nd * R.L1 + f.l, q = (f.j * nmax + f.n) * R.L1 + f.l;
        v += rv * Bt[a] * U[q] + B[a] * V[q];
    }
    return v;
}

__device__ __forceinline__ double cc_term(const double* rec, const RecLayout& R, int nmax, int s, int d, int f) {
    const int per_s = nmax * R.L2;
    const int fs = f / per_s;
    if (fs != s) return 0.0;
    const int rem = f - fs * per_s;
    const int nn = rem / R.L2, lmp = rem - nn * R.L2;
    const int lm = cc_src(lmp), l = lm_l(lm);
    const double* Y = rec + R.offU;
    const double* dY = rec + R.offV + (size_t)d * R.L2;
    const double v = rec[d] * rec[R.offBt + nn * R.L1 + l] * Y[lm] + rec[R.offB + nn * R.L1 + l] * dY[lm];
    return lmp >= 4 ? v * 0.17958712212516656 /* pi^-3/2 */ : v;
}

template <typename TO>
__global__ __launch_bounds__(256) void dense_kernel(cusoap_basis_t b, int mode, int n_atoms,
                                                    const int32_t* __restrict__ atom_sp,
                                                    const int32_t* __restrict__ center_attach,
                                                    const int64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ inv_atom, int n_inc,
                                                    const uint32_t* __restrict__ ft, int n_feat,
                                                    const double* __restrict__ ws, TO* __restrict__ out,
                                                    int32_t* __restrict__ err) {
    __shared__ unsigned long long keys[kJacCap];
    __shared__ int nkeys;
    __shared__ double prefs[kMaxL + 2];
    const int n = b.n_max, S = b.n_species;
    const RecLayout R = rec_layout(S, n, b.l_max, mode);
    const int c = blockIdx.x;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    if (threadIdx.x == 0) nkeys = 0;
    if ((int)threadIdx.x <= b.l_max) prefs[threadIdx.x] = sqrt(8.0 * M_PI * M_PI / (2 * threadIdx.x + 1));
    __syncthreads();
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int a = nbr[e] % n_atoms;
        const int row = inv_atom[a];
        if (row >= 0) {
            const int slot = atomicAdd(&nkeys, 1);
            if (slot < kJacCap) keys[slot] = ((unsigned long long)row << 32) | (unsigned long long)(e - e0);
        }
    }
    __syncthreads();
    const int k = nkeys;
    if (k > kJacCap) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    int p2 = 1;
    while (p2 < k) p2 <<= 1;
    for (int i = k + threadIdx.x; i < p2; i += blockDim.x) keys[i] = ~0ull;
    __syncthreads();
    for (int kk = 2; kk <= p2; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < p2; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long x = keys[i], y = keys[ixj];
                    if ((x > y) == ((i & kk) == 0)) {
                        keys[i] = y;
                        keys[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    int att_row = -1;
    if (center_attach) {
        const int aa = center_attach[c];
        if (aa >= 0) att_row = inv_atom[aa];
    }
    const size_t row_len = 3 * (size_t)n_feat;
    TO* outc = out + (size_t)c * n_inc * row_len;
    int row = 0, i = 0;
    while (row < n_inc) {
        int nxt = i < k ? (int)(keys[i] >> 32) : n_inc;
        if (att_row >= row && att_row < nxt) nxt = att_row;
        zero_fill(outc + (size_t)row * row_len, (size_t)(nxt - row) * row_len);
        if (nxt >= n_inc) break;
        int j = i;
        while (j < k && (int)(keys[j] >> 32) == nxt) ++j;
        TO* dst = outc + (size_t)nxt * row_len;
        for (int t = threadIdx.x; t < (int)row_len; t += blockDim.x) {
            const int d = t / n_feat, f = t - d * n_feat;
            double v = 0.0;
            if (mode == CUSOAP_MODE_PS) {
                const FeatDec fd = dec(ft[f]);
                for (int q = i; q < j; ++q) {
                    const int64_t e = e0 + (int64_t)(keys[q] & 0xffffffffull);
                    v += ps_term(ws + (size_t)e * R.stride, R, n, atom_sp[nbr[e] % n_atoms], d, fd);
                }
                if (nxt == att_row)
                    for (int64_t e = e0; e < e1; ++e)
                        v -= ps_term(ws + (size_t)e * R.stride, R, n, atom_sp[nbr[e] % n_atoms], d, fd);
                v *= prefs[fd.l];
            } else {
                for (int q = i; q < j; ++q) {
                    const int64_t e = e0 + (int64_t)(keys[q] & 0xffffffffull);
                    v += cc_term(ws + (size_t)e * R.stride, R, n, atom_sp[nbr[e] % n_atoms], d, f);
                }
                if (nxt == att_row)
                    for (int64_t e = e0; e < e1; ++e)
                        v -= cc_term(ws + (size_t)e * R.stride, R, n, atom_sp[nbr[e] % n_atoms], d, f);
            }
            dst[t] = (TO)v;
        }
        row = nxt + 1;
        i = j;
    }
}

template <int LMAX>
int launch_edge(const cusoap_basis_t& b, int mode, const double* pos, int n_atoms, const double* shifts,
                const double* centers, int n_centers, const int64_t* offsets, const int32_t* nbr,
                const double* coeffs, double* ws, hipStream_t st) {
    const RecLayout R = rec_layout(b.n_species, b.n_max, LMAX, mode);
    const size_t smem = sizeof(double) * 4 * (size_t)(4 * R.L2 + 4 * R.NL);
    prof_begin("jacobian_edges", st);
    if (b.rbf == CUSOAP_RBF_GTO)
        edge_kernel<LMAX, CUSOAP_RBF_GTO><<<n_centers, 256, smem, st>>>(b, mode, pos, n_atoms, shifts, centers,
                                                                        offsets, nbr, coeffs, ws);
    else
        edge_kernel<LMAX, CUSOAP_RBF_POLY><<<n_centers, 256, smem, st>>>(b, mode, pos, n_atoms, shifts, centers,
                                                                         offsets, nbr, coeffs, ws);
    prof_end("jacobian_edges", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace

size_t jacobian_ws_bytes(const cusoap_basis_t* b, int mode, int64_t n_edges) {
    const RecLayout R = rec_layout(b->n_species, b->n_max, b->l_max, mode);
    return align_up(sizeof(double) * (size_t)R.stride * (size_t)std::max<int64_t>(n_edges, 1), 256) + 256;
}

int jacobian_dense_impl(const cusoap_basis_t* b, int mode, const double* pos, int n_atoms, const int32_t* atom_sp,
                        const double* shifts, int n_img, const double* centers, int n_centers,
                        const int32_t* center_attach, const int64_t* offsets, const int32_t* nbr, int64_t n_edges,
                        const double* coeffs, const int32_t* inv_atom, int n_inc, const uint32_t* ft, int n_feat,
                        void* ws, size_t ws_bytes, void* out, int odt, hipStream_t st) {
    (void)n_img;
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: n_max must be 1..16 and l_max 0..9");
    if (mode != CUSOAP_MODE_PS && mode != CUSOAP_MODE_CC) return fail(CUSOAP_ERR_ARG, "jacobian: unknown mode");
    if (b->n_species > 255 || b->n_max > 63) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: feature table range");
    const size_t need = jacobian_ws_bytes(b, mode, n_edges);
    if (ws_bytes < need) return fail(CUSOAP_ERR_WORKSPACE, "jacobian: workspace too small");
    if (n_centers == 0 || n_inc == 0 || n_feat == 0) return CUSOAP_OK;
    double* rec = (double*)ws;
    int32_t* err = (int32_t*)((char*)ws + need - 256);
    CUSOAP_HIP(hipMemsetAsync(err, 0, sizeof(int32_t), st));
    int rc = CUSOAP_OK;
#define CUSOAP_CASE(L) \
    case L: rc = launch_edge<L>(*b, mode, pos, n_atoms, shifts, centers, n_centers, offsets, nbr, coeffs, rec, st); break;
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: l_max out of range");
    }
#undef CUSOAP_CASE
    if (rc) return rc;
    prof_begin("jacobian_dense", st);
    if (odt == CUSOAP_DT_F32)
        dense_kernel<float><<<n_centers, 256, 0, st>>>(*b, mode, n_atoms, atom_sp, center_attach, offsets, nbr,
                                                       inv_atom, n_inc, ft, n_feat, rec, (float*)out, err);
    else
        dense_kernel<double><<<n_centers, 256, 0, st>>>(*b, mode, n_atoms, atom_sp, center_attach, offsets, nbr,
                                                        inv_atom, n_inc, ft, n_feat, rec, (double*)out, err);
    prof_end("jacobian_dense", st);
    CUSOAP_LAUNCH_CHECK();
    int32_t herr = 0;
    CUSOAP_HIP(hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CUSOAP_HIP(hipStreamSynchronize(st));
    if (herr) return fail(CUSOAP_ERR_CAPACITY, "jacobian: a center has more than 4096 included edges");
    return CUSOAP_OK;
}

}  // namespace cusoap
