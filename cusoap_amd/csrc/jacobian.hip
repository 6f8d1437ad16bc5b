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
// Three kernels:
//  rowkey_kernel   one workgroup per center: the center's included edges as
//                  (output row, edge) keys, bitonic-sorted in LDS -> global.
//  records_kernel  one workgroup per center, tiles of up to 64 edges: solid
//                  harmonics with gradients (thread per edge), radial factors
//                  (thread per (edge, k, l)), then Bw/Bt and the U/V
//                  contractions (thread per output) into one float64 record
//                  per edge:  rvec[3] | Bw[n][l] | Bt[n][l] | ps: U[Sn][l], V[3][Sn][l]
//                                                              cc: Y[L2], dY[3][L2]
//  dense_kernel    one workgroup per (center, 64-row chunk); each wavefront
//                  takes one output row at a time: rows without an edge are
//                  stored as 16-byte non-temporal zeros, rows with one edge are
//                  computed from the record staged in LDS and stored 4 elements
//                  per lane.  Every output element is stored exactly once.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kJacCap = 4096;   // max included edges per center (row-key sort in LDS)
constexpr int kRowsPerBlock = 64;

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

// wavefront-local LDS hand-off (all lanes of one wavefront)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// row keys
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rowkey_kernel(int n_atoms, const int64_t* __restrict__ offsets,
                                                     const int32_t* __restrict__ nbr,
                                                     const int32_t* __restrict__ inv_atom,
                                                     unsigned long long* __restrict__ rowkeys,
                                                     int32_t* __restrict__ kcount, int32_t* __restrict__ err) {
    __shared__ unsigned long long keys[kJacCap];
    __shared__ int nkeys;
    const int c = blockIdx.x;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    if (threadIdx.x == 0) nkeys = 0;
    __syncthreads();
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int row = inv_atom[nbr[e] % n_atoms];
        if (row >= 0) {
            const int slot = atomicAdd(&nkeys, 1);
            if (slot < kJacCap) keys[slot] = ((unsigned long long)row << 32) | (unsigned long long)(e - e0);
        }
    }
    __syncthreads();
    const int k = nkeys;
    if (k > kJacCap) {
        if (threadIdx.x == 0) {
            atomicOr(err, 1);
            kcount[c] = 0;
        }
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
    for (int i = threadIdx.x; i < k; i += blockDim.x) rowkeys[e0 + i] = keys[i];
    if (threadIdx.x == 0) kcount[c] = k;
}

// ---------------------------------------------------------------------------
// per-edge records
// ---------------------------------------------------------------------------
template <int LMAX, int RBF>
__global__ __launch_bounds__(256) void records_kernel(cusoap_basis_t b, int mode, int TE,
                                                      const double* __restrict__ pos, int n_atoms,
                                                      const double* __restrict__ shifts,
                                                      const double* __restrict__ centers,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ nbr,
                                                      const double* __restrict__ coeffs, double* __restrict__ rec) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const int n = b.n_max, S = b.n_species, NL = n * L1;
    const RecLayout R = rec_layout(S, n, LMAX, mode);
    extern __shared__ __align__(16) unsigned char smem[];
    double* Cs = (double*)smem;            // [S n L2]
    double* Ys = Cs + S * n * L2;          // [TE][4 L2]
    double* Qs = Ys + TE * 4 * L2;         // [TE][NL]  GTO primitives / poly B
    double* Qh = Qs + TE * NL;             // [TE][NL]  poly Bhat
    double* geo = Qh + TE * NL;            // [TE][8]   rx ry rz r2 r w dw/r
    const int c = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    if (mode == CUSOAP_MODE_PS)
        for (int i = tid; i < S * n * L2; i += nth) Cs[i] = coeffs[(size_t)c * S * n * L2 + i];
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    for (int64_t t0 = e0; t0 < e1; t0 += TE) {
        const int nt = (int)((e1 - t0) < TE ? (e1 - t0) : TE);
        // ---- geometry
        if (tid < nt) {
            const int64_t e = t0 + tid;
            const int ext = nbr[e];
            const int img = ext / n_atoms, a = ext - img * n_atoms;
            const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
            const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
            const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            const double r = sqrt(r2);
            double w = 1.0, dw = 0.0;
            if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
            double* g = geo + tid * 8;
            g[0] = rx; g[1] = ry; g[2] = rz; g[3] = r2; g[4] = r; g[5] = w; g[6] = r > 0.0 ? dw / r : 0.0;
            double* rp = rec + (size_t)e * R.stride;
            rp[0] = rx; rp[1] = ry; rp[2] = rz;
        }
        __syncthreads();
        // ---- solid harmonics (one item per edge) and radial factors
        const int n_rad = (RBF == CUSOAP_RBF_GTO) ? nt * NL : nt * n;
        for (int item = tid; item < nt + n_rad; item += nth) {
            if (item < nt) {
                const double* g = geo + item * 8;
                double* y = Ys + item * 4 * L2;
                solid_harmonics<LMAX, true, double>(g[0], g[1], g[2], [&](int lm, double v, double gx, double gy, double gz) {
                    y[lm] = v;
                    y[L2 + lm] = gx;
                    y[2 * L2 + lm] = gy;
                    y[3 * L2 + lm] = gz;
                });
            } else if (RBF == CUSOAP_RBF_GTO) {
                const int it = item - nt;
                const int e = it / NL, t = it - e * NL;
                const int k = t / L1, l = t - k * L1;
                Qs[e * NL + t] = b.gto_pref[l * n + k] * exp(-b.gto_gamma[l * n + k] * geo[e * 8 + 3]);
            } else {
                const int it = item - nt;
                const int e = it / n, k = it - e * n;
                const double r = geo[e * 8 + 4];
                double* qb = Qs + e * NL + k * L1;
                double* qh = Qh + e * NL + k * L1;
                if (r < 1e-4) {
#pragma unroll
                    for (int l = 0; l < L1; ++l) {
                        qb[l] = b.poly_B0[k * L1 + l];
                        qh[l] = b.poly_Bh0[k * L1 + l];
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
                        qb[l] = I[l] / rl;
                        qh[l] = (2.0 * b.eta) * (J[l] - r * I[l]) / (rl * r);
                        rl *= r;
                    }
                }
            }
        }
        __syncthreads();
        // ---- Bw / Bt and the U, V contractions (or Y, dY for cc)
        const int per_uv = (mode == CUSOAP_MODE_PS) ? 4 * R.SNL : 4 * L2;
        for (int item = tid; item < nt * (NL + per_uv); item += nth) {
            if (item < nt * NL) {
                const int e = item / NL, t = item - e * NL;
                const int nn = t / L1, l = t - nn * L1;
                double B, Bh;
                if (RBF == CUSOAP_RBF_GTO) {
                    const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
                    B = 0.0;
                    Bh = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double q = Qs[e * NL + k * L1 + l];
                        B += beta[k] * q;
                        Bh += beta[k] * (-2.0 * b.gto_gamma[l * n + k]) * q;
                    }
                } else {
                    B = Qs[e * NL + t];
                    Bh = Qh[e * NL + t];
                }
                const double* g = geo + e * 8;
                double* rp = rec + (size_t)(t0 + e) * R.stride;
                rp[R.offB + t] = g[5] * B;
                rp[R.offBt + t] = g[5] * Bh + g[6] * B;
            } else {
                const int it = item - nt * NL;
                const int e = it / per_uv, t = it - e * per_uv;
                double* rp = rec + (size_t)(t0 + e) * R.stride;
                const double* y = Ys + e * 4 * L2;
                if (mode == CUSOAP_MODE_PS) {
                    const int comp = t / R.SNL, rem = t - comp * R.SNL;
                    const int kq = rem / L1, l = rem - kq * L1;
                    const double* yv = y + comp * L2 + l * l;
                    const double* cv = Cs + kq * L2 + l * l;
                    double acc = 0.0;
                    for (int m = 0; m < 2 * l + 1; ++m) acc += yv[m] * cv[m];
                    rp[R.offU + t] = acc;
                } else {
                    rp[R.offU + t] = y[t];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// dense rows
// ---------------------------------------------------------------------------
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
    const double* V = rec + R.offV + d * R.SNL;
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
    const double v = rec[d] * rec[R.offBt + nn * R.L1 + l] * rec[R.offU + lm] +
                     rec[R.offB + nn * R.L1 + l] * rec[R.offV + d * R.L2 + lm];
    return lmp >= 4 ? v * 0.17958712212516656 /* pi^-3/2 */ : v;
}

__device__ __forceinline__ double elem_value(const double* rp, const RecLayout& R, int mode, int nmax, int s, int t,
                                             int n_feat, const uint32_t* fts, const double* prefs) {
    const int d = t / n_feat, f = t - d * n_feat;
    if (mode == CUSOAP_MODE_PS) {
        const FeatDec fd = dec(fts[f]);
        return prefs[fd.l] * ps_term(rp, R, nmax, s, d, fd);
    }
    return cc_term(rp, R, nmax, s, d, f);
}

template <typename TO>
__device__ __forceinline__ void store4(TO* p, double a, double b, double c, double d);
template <>
__device__ __forceinline__ void store4<float>(float* p, double a, double b, double c, double d) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const f32x4 v = {(float)a, (float)b, (float)c, (float)d};
    __builtin_nontemporal_store(v, (f32x4*)p);
}
template <>
__device__ __forceinline__ void store4<double>(double* p, double a, double b, double c, double d) {
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const f64x2 v0 = {a, b}, v1 = {c, d};
    __builtin_nontemporal_store(v0, (f64x2*)p);
    __builtin_nontemporal_store(v1, (f64x2*)(p + 2));
}

// lower bound of `row` in keys[lo, hi) (keys sorted; row in the upper 32 bits)
__device__ __forceinline__ int lower_row(const unsigned long long* keys, int lo, int hi, int row) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)(keys[mid] >> 32) < row) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// VEC: the row length is a multiple of 4 elements and the output is 16-byte aligned
template <typename TO, bool VEC>
__global__ __launch_bounds__(256) void dense_kernel(cusoap_basis_t b, int mode, int n_atoms,
                                                    const int32_t* __restrict__ atom_sp,
                                                    const int32_t* __restrict__ center_attach,
                                                    const int64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ inv_atom,
                                                    const unsigned long long* __restrict__ rowkeys,
                                                    const int32_t* __restrict__ kcount, int n_inc, int chunks,
                                                    const uint32_t* __restrict__ ft, int n_feat,
                                                    const double* __restrict__ rec, TO* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = b.n_max, S = b.n_species;
    const RecLayout R = rec_layout(S, n, b.l_max, mode);
    double* prefs = (double*)smem;                          // [16]
    double* recs = prefs + 16;                              // [4][stride]
    uint32_t* fts = (uint32_t*)(recs + 4 * R.stride);       // [n_feat]
    const int c = blockIdx.x / chunks, ch = blockIdx.x - c * chunks;
    const int r0 = ch * kRowsPerBlock;
    const int r1 = (r0 + kRowsPerBlock < n_inc) ? r0 + kRowsPerBlock : n_inc;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int kc = kcount[c];
    const unsigned long long* keys = rowkeys + e0;
    if (mode == CUSOAP_MODE_PS)
        for (int i = threadIdx.x; i < n_feat; i += blockDim.x) fts[i] = ft[i];
    if ((int)threadIdx.x <= b.l_max) prefs[threadIdx.x] = sqrt(8.0 * M_PI * M_PI / (2 * threadIdx.x + 1));
    int att_row = -1;
    if (center_attach) {
        const int aa = center_attach[c];
        if (aa >= 0) att_row = inv_atom[aa];
    }
    __syncthreads();
    const int lo = lower_row(keys, 0, kc, r0);
    const int hi = lower_row(keys, lo, kc, r1);
    const size_t row_len = 3 * (size_t)n_feat;
    double* myrec = recs + wv * R.stride;
    for (int r = r0 + wv; r < r1; r += 4) {
        TO* dst = out + ((size_t)c * n_inc + r) * row_len;
        const int q0 = lower_row(keys, lo, hi, r);
        const int q1 = lower_row(keys, q0, hi, r + 1);
        if (q0 == q1 && r != att_row) {
            if (VEC) {
                typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
                const u32x4 z = {0u, 0u, 0u, 0u};
                u32x4* p = (u32x4*)dst;
                const size_t nv = row_len * sizeof(TO) / 16;
                for (size_t i = lane; i < nv; i += 64) __builtin_nontemporal_store(z, &p[i]);
            } else {
                for (size_t i = lane; i < row_len; i += 64) dst[i] = TO(0);
            }
            continue;
        }
        if (q1 - q0 == 1 && r != att_row) {
            // one edge: stage its record in LDS (wavefront-private slot)
            const int64_t e = e0 + (int64_t)(keys[q0] & 0xffffffffull);
            const double* rp = rec + (size_t)e * R.stride;
            for (int i = lane; i < R.stride; i += 64) myrec[i] = rp[i];
            wave_sync();
            const int s = atom_sp[nbr[e] % n_atoms];
            if (VEC) {
                for (size_t t = 4 * (size_t)lane; t < row_len; t += 256) {
                    const double v0 = elem_value(myrec, R, mode, n, s, (int)t, n_feat, fts, prefs);
                    const double v1 = elem_value(myrec, R, mode, n, s, (int)t + 1, n_feat, fts, prefs);
                    const double v2 = elem_value(myrec, R, mode, n, s, (int)t + 2, n_feat, fts, prefs);
                    const double v3 = elem_value(myrec, R, mode, n, s, (int)t + 3, n_feat, fts, prefs);
                    store4<TO>(dst + t, v0, v1, v2, v3);
                }
            } else {
                for (size_t t = lane; t < row_len; t += 64)
                    dst[t] = (TO)elem_value(myrec, R, mode, n, s, (int)t, n_feat, fts, prefs);
            }
            wave_sync();
            continue;
        }
        // several edges (periodic images) and/or the attach row: read records from global
        for (size_t t = lane; t < row_len; t += 64) {
            double v = 0.0;
            for (int q = q0; q < q1; ++q) {
                const int64_t e = e0 + (int64_t)(keys[q] & 0xffffffffull);
                v += elem_value(rec + (size_t)e * R.stride, R, mode, n, atom_sp[nbr[e] % n_atoms], (int)t, n_feat,
                                fts, prefs);
            }
            if (r == att_row)
                for (int64_t e = e0; e < e1; ++e)
                    v -= elem_value(rec + (size_t)e * R.stride, R, mode, n, atom_sp[nbr[e] % n_atoms], (int)t,
                                    n_feat, fts, prefs);
            dst[t] = (TO)v;
        }
    }
}

int records_tile(const RecLayout& R, int S, int n, size_t* smem) {
    const size_t fixed = (size_t)S * n * R.L2;
    const size_t per = 4 * (size_t)R.L2 + 2 * (size_t)R.NL + 8;
    size_t room = 8192 > fixed ? 8192 - fixed : 0;     // doubles in 64 KiB
    size_t te = room / per;
    if (te > 64) te = 64;
    if (te < 8) te = 8;
    *smem = sizeof(double) * (fixed + te * per);
    return (int)te;
}

template <int LMAX>
int launch_records(const cusoap_basis_t& b, int mode, const double* pos, int n_atoms, const double* shifts,
                   const double* centers, int n_centers, const int64_t* offsets, const int32_t* nbr,
                   const double* coeffs, double* rec, hipStream_t st) {
    const RecLayout R = rec_layout(b.n_species, b.n_max, LMAX, mode);
    size_t smem = 0;
    const int TE = records_tile(R, b.n_species, b.n_max, &smem);
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: coefficient block too large for LDS");
    prof_begin("jacobian_edges", st);
    if (b.rbf == CUSOAP_RBF_GTO) {
        auto kern = records_kernel<LMAX, CUSOAP_RBF_GTO>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        kern<<<n_centers, 256, smem, st>>>(b, mode, TE, pos, n_atoms, shifts, centers, offsets, nbr, coeffs, rec);
    } else {
        auto kern = records_kernel<LMAX, CUSOAP_RBF_POLY>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        kern<<<n_centers, 256, smem, st>>>(b, mode, TE, pos, n_atoms, shifts, centers, offsets, nbr, coeffs, rec);
    }
    prof_end("jacobian_edges", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <typename TO, bool VEC>
int launch_dense(const cusoap_basis_t& b, int mode, int n_atoms, const int32_t* atom_sp, const int32_t* center_attach,
                 const int64_t* offsets, const int32_t* nbr, const int32_t* inv_atom,
                 const unsigned long long* rowkeys, const int32_t* kcount, int n_centers, int n_inc,
                 const uint32_t* ft, int n_feat, const double* rec, void* out, hipStream_t st) {
    const RecLayout R = rec_layout(b.n_species, b.n_max, b.l_max, mode);
    const size_t smem = sizeof(double) * (16 + 4 * (size_t)R.stride) + sizeof(uint32_t) * (size_t)n_feat;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: record too large for LDS");
    const int chunks = (n_inc + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t grid = (int64_t)n_centers * chunks;
    if (grid >= ((int64_t)1 << 31)) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: output too large");
    auto kern = dense_kernel<TO, VEC>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("jacobian_dense", st);
    kern<<<(unsigned)grid, 256, smem, st>>>(b, mode, n_atoms, atom_sp, center_attach, offsets, nbr, inv_atom, rowkeys,
                                            kcount, n_inc, chunks, ft, n_feat, rec, (TO*)out);
    prof_end("jacobian_dense", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

struct JacWs {
    double* rec;
    unsigned long long* rowkeys;
    int32_t* kcount;
    int32_t* err;
    size_t bytes;
};

JacWs carve_jac(void* base, const RecLayout& R, int64_t n_edges, int n_centers) {
    JacWs w;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return (char*)base + o;
    };
    const int64_t E = std::max<int64_t>(n_edges, 1);
    w.rec = (double*)take(sizeof(double) * (size_t)R.stride * (size_t)E);
    w.rowkeys = (unsigned long long*)take(sizeof(unsigned long long) * (size_t)E);
    w.kcount = (int32_t*)take(sizeof(int32_t) * (size_t)std::max(n_centers, 1));
    w.err = (int32_t*)take(sizeof(int32_t) * 4);
    w.bytes = off;
    return w;
}

}  // namespace

size_t jacobian_ws_bytes(const cusoap_basis_t* b, int mode, int64_t n_edges, int n_centers) {
    const RecLayout R = rec_layout(b->n_species, b->n_max, b->l_max, mode);
    return carve_jac(nullptr, R, n_edges, n_centers).bytes;
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
    const RecLayout R = rec_layout(b->n_species, b->n_max, b->l_max, mode);
    JacWs w = carve_jac(ws, R, n_edges, n_centers);
    if (ws_bytes < w.bytes) return fail(CUSOAP_ERR_WORKSPACE, "jacobian: workspace too small");
    if (n_centers == 0 || n_inc == 0 || n_feat == 0) return CUSOAP_OK;
    CUSOAP_HIP(hipMemsetAsync(w.err, 0, sizeof(int32_t), st));
    rowkey_kernel<<<n_centers, 256, 0, st>>>(n_atoms, offsets, nbr, inv_atom, w.rowkeys, w.kcount, w.err);
    CUSOAP_LAUNCH_CHECK();
    int rc = CUSOAP_OK;
#define CUSOAP_CASE(L) \
    case L: rc = launch_records<L>(*b, mode, pos, n_atoms, shifts, centers, n_centers, offsets, nbr, coeffs, w.rec, st); break;
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: l_max out of range");
    }
#undef CUSOAP_CASE
    if (rc) return rc;
    // 4-element groups per lane: every row length must be a multiple of 4 elements
    const bool vec = ((3 * (size_t)n_feat) % 4 == 0) && (((uintptr_t)out & 15) == 0);
#define CUSOAP_DENSE(TO, V) \
    launch_dense<TO, V>(*b, mode, n_atoms, atom_sp, center_attach, offsets, nbr, inv_atom, w.rowkeys, w.kcount, \
                        n_centers, n_inc, ft, n_feat, w.rec, out, st)
    if (odt == CUSOAP_DT_F32) rc = vec ? CUSOAP_DENSE(float, true) : CUSOAP_DENSE(float, false);
    else rc = vec ? CUSOAP_DENSE(double, true) : CUSOAP_DENSE(double, false);
#undef CUSOAP_DENSE
    if (rc) return rc;
    int32_t herr = 0;
    CUSOAP_HIP(hipMemcpyAsync(&herr, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CUSOAP_HIP(hipStreamSynchronize(st));
    if (herr) return fail(CUSOAP_ERR_CAPACITY, "jacobian: a center has more than 4096 included edges");
    return CUSOAP_OK;
}

}  // namespace cusoap
