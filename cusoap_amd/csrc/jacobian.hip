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
// Kernels:
//  rowkey_kernel   one workgroup per center: the center's included edges as
//                  (output row, edge) keys, bitonic-sorted in LDS -> global.
//  records_kernel  workgroup (center, y) takes the center's edge tiles y, y + gy, ...
//                  (up to 32 edges each; gy from the mean tiles per center): solid
//                  harmonics with gradients (thread per edge), radial factors
//                  (thread per (edge, k, l)), then Bw/Bt and the U/V
//                  contractions (thread per output) into one record per edge,
//                  float32 when the rows are computed in float32, else float64:
//                      rvec[3] | Bw[n][l] | Bt[n][l] | ps: U[Sn][l], V[3][Sn][l]
//                                                      cc: Y[L2], dY[3][L2]
//  dense_fused_kernel  (one image, no attached center, GTO, power spectrum: one edge
//                  per output row) forms each row's record in LDS in the wavefront that
//                  stores the row, instead of records_kernel + dense_kernel.
//  dense_kernel    one workgroup per (center, 32-row chunk): each run of
//                  consecutive rows without an edge is stored by the whole
//                  workgroup as one flat range of 16-byte zeros; then each
//                  wavefront takes one row with edges at a time, computed from
//                  the record staged in LDS, 4 elements per lane.  Every output
//                  element is stored exactly once.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kJacCap = 4096;   // max included edges per center (row-key sort in LDS)
// 32-row chunks (161 KB of float32 output per workgroup at 420 features): the output
// region the resident workgroups write at one time is half that of 64-row chunks, which
// measured ~1.5% faster on (H2O)_1000 (9.09 against 9.24 ms, means over 18 buffers)
constexpr int kRowsPerBlock = 32;
constexpr int kChunkKeys = 512;  // keys of one chunk staged in LDS (more: read from global)

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
                                                     int32_t* __restrict__ kcount, int chunks,
                                                     int32_t* __restrict__ chunk_lo, int32_t* __restrict__ err) {
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
        for (int ch = threadIdx.x; ch <= chunks; ch += blockDim.x) chunk_lo[(size_t)c * (chunks + 1) + ch] = 0;
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
    // first key of every row chunk (chunk_lo[chunks] = k)
    for (int ch = threadIdx.x; ch <= chunks; ch += blockDim.x) {
        const int row = ch * kRowsPerBlock;
        int lo = 0, hi = k;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)(keys[mid] >> 32) < row) lo = mid + 1;
            else hi = mid;
        }
        chunk_lo[(size_t)c * (chunks + 1) + ch] = lo;
    }
    if (threadIdx.x == 0) kcount[c] = k;
}

// ---------------------------------------------------------------------------
// per-edge records
// ---------------------------------------------------------------------------
template <int LMAX, int RBF, typename TR>
__global__ __launch_bounds__(256) void records_kernel(cusoap_basis_t b, int mode, int TE,
                                                      const double* __restrict__ pos, int n_atoms,
                                                      const double* __restrict__ shifts,
                                                      const double* __restrict__ centers,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ nbr,
                                                      const double* __restrict__ coeffs, TR* __restrict__ rec) {
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
    for (int64_t t0 = e0 + (int64_t)blockIdx.y * TE; t0 < e1; t0 += (int64_t)gridDim.y * TE) {
        const int nt = (int)((e1 - t0) < TE ? (e1 - t0) : TE);
        // ---- geometry
        if (tid < nt) {
            const int64_t e = t0 + tid;
            const int ext = nbr[e];
            int img, a;
            split_ext(ext, n_atoms, 1.0 / (double)n_atoms, img, a);
            const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
            const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
            const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            const double r = sqrt(r2);
            double w = 1.0, dw = 0.0;
            if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
            double* g = geo + tid * 8;
            g[0] = rx; g[1] = ry; g[2] = rz; g[3] = r2; g[4] = r; g[5] = w; g[6] = r > 0.0 ? dw / r : 0.0;
            TR* rp = rec + (size_t)e * R.stride;
            rp[0] = (TR)rx; rp[1] = (TR)ry; rp[2] = (TR)rz;
        }
        __syncthreads();
        // ---- solid harmonics (thread per edge) and radial factors.  The GTO primitives
        // use a fixed (k, l) slot per thread, decoded once, looping over the tile's edges.
        if (tid < nt) {
            const double* g = geo + tid * 8;
            double* y = Ys + tid * 4 * L2;
            solid_harmonics<LMAX, true, double>(g[0], g[1], g[2], [&](int lm, double v, double gx, double gy, double gz) {
                y[lm] = v;
                y[L2 + lm] = gx;
                y[2 * L2 + lm] = gy;
                y[3 * L2 + lm] = gz;
            });
        }
        if (RBF == CUSOAP_RBF_GTO) {
            const int per_slot = nth / NL;      // threads per (k, l) slot (NL <= 160 < nth)
            const int slot = tid % NL, e_first = tid / NL;
            if (e_first < per_slot) {
                const int k = slot / L1, l = slot - k * L1;
                const double pref = b.gto_pref[l * n + k], gam = b.gto_gamma[l * n + k];
                for (int e = e_first; e < nt; e += per_slot) Qs[e * NL + slot] = pref * exp(-gam * geo[e * 8 + 3]);
            }
        } else {
            for (int it = tid; it < nt * n; it += nth) {
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
        // ---- Bw / Bt: a fixed (n, l) slot per thread, decoded once, looping over the edges
        {
            const int per_slot = nth / NL;
            const int t = tid % NL, e_first = tid / NL;
            if (e_first < per_slot) {
                const int nn = t / L1, l = t - nn * L1;
                const double* beta = (RBF == CUSOAP_RBF_GTO) ? b.gto_beta + ((size_t)l * n + nn) * n : nullptr;
                for (int e = e_first; e < nt; e += per_slot) {
                    double B, Bh;
                    if (RBF == CUSOAP_RBF_GTO) {
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
                    TR* rp = rec + (size_t)(t0 + e) * R.stride;
                    rp[R.offB + t] = (TR)(g[5] * B);
                    rp[R.offBt + t] = (TR)(g[5] * Bh + g[6] * B);
                }
            }
        }
        // ---- U, V contractions (or Y, dY for cc): thread per output slot t, decoded once
        // and looping over the tile's edges, with the coefficient row in registers
        const int per_uv = (mode == CUSOAP_MODE_PS) ? 4 * R.SNL : 4 * L2;
        for (int t = tid; t < per_uv; t += nth) {
            TR* rp = rec + (size_t)t0 * R.stride + R.offU + t;
            if (mode == CUSOAP_MODE_PS) {
                const int comp = t / R.SNL, rem = t - comp * R.SNL;
                const int kq = rem / L1, l = rem - kq * L1;
                const int nm = 2 * l + 1;
                double cr[2 * LMAX + 1];
#pragma unroll
                for (int m = 0; m < 2 * LMAX + 1; ++m) cr[m] = m < nm ? Cs[kq * L2 + l * l + m] : 0.0;
                const double* yv = Ys + comp * L2 + l * l;
                for (int e = 0; e < nt; ++e) {
                    double acc = 0.0;
#pragma unroll
                    for (int m = 0; m < 2 * LMAX + 1; ++m)
                        if (m < nm) acc += yv[e * 4 * L2 + m] * cr[m];
                    rp[(size_t)e * R.stride] = (TR)acc;
                }
            } else {
                for (int e = 0; e < nt; ++e) rp[(size_t)e * R.stride] = (TR)Ys[e * 4 * L2 + t];
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

template <typename TR>
__device__ __forceinline__ double ps_term(const TR* rec, const RecLayout& R, int nmax, int s, int d,
                                          const FeatDec& f) {
    const double rv = rec[d];
    const TR* B = rec + R.offB;
    const TR* Bt = rec + R.offBt;
    const TR* U = rec + R.offU;
    const TR* V = rec + R.offV + d * R.SNL;
    double v = 0.0;
    if (f.j == s) {
        const int a = f.n * R.L1 + f.l, q = (f.jd * nmax + f.nd) * R.L1 + f.l;
        v += rv * (double)Bt[a] * (double)U[q] + (double)B[a] * (double)V[q];
    }
    if (f.jd == s) {
        const int a = f.nd * R.L1 + f.l, q = (f.j * nmax + f.n) * R.L1 + f.l;
        v += rv * (double)Bt[a] * (double)U[q] + (double)B[a] * (double)V[q];
    }
    return v;
}

template <typename TR>
__device__ __forceinline__ double cc_term(const TR* rec, const RecLayout& R, int nmax, int s, int d, int f) {
    const int per_s = nmax * R.L2;
    const int fs = f / per_s;
    if (fs != s) return 0.0;
    const int rem = f - fs * per_s;
    const int nn = rem / R.L2, lmp = rem - nn * R.L2;
    const int lm = cc_src(lmp), l = lm_l(lm);
    const double v = (double)rec[d] * (double)rec[R.offBt + nn * R.L1 + l] * (double)rec[R.offU + lm] +
                     (double)rec[R.offB + nn * R.L1 + l] * (double)rec[R.offV + d * R.L2 + lm];
    return lmp >= 4 ? v * 0.17958712212516656 /* pi^-3/2 */ : v;
}

template <typename TR>
__device__ __forceinline__ double elem_value(const TR* rp, const RecLayout& R, int mode, int nmax, int s, int t,
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
    *(f32x4*)p = v;
}
template <>
__device__ __forceinline__ void store4<double>(double* p, double a, double b, double c, double d) {
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const f64x2 v0 = {a, b}, v1 = {c, d};
    *(f64x2*)p = v0;
    *(f64x2*)(p + 2) = v1;
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

// Feature plan of the power spectrum, per feature f (decoded once per workgroup):
// species j, jd | l | record offsets a1 = n l, a2 = nd l (into Bw/Bt) and
// q1 = (jd nd) l, q2 = (j n) l (into U/V).
struct FPlan {
    int j, jd, l, a1, a2, q1, q2;
};

__device__ __forceinline__ unsigned long long pack_plan(int j, int jd, int l, int a1, int a2, int q1, int q2) {
    return (unsigned long long)j | ((unsigned long long)jd << 8) | ((unsigned long long)l << 16) |
           ((unsigned long long)a1 << 20) | ((unsigned long long)a2 << 30) | ((unsigned long long)q1 << 40) |
           ((unsigned long long)q2 << 52);
}

__device__ __forceinline__ FPlan unpack_plan(unsigned long long v) {
    FPlan p;
    p.j = (int)(v & 0xff);
    p.jd = (int)((v >> 8) & 0xff);
    p.l = (int)((v >> 16) & 0xf);
    p.a1 = (int)((v >> 20) & 0x3ff);
    p.a2 = (int)((v >> 30) & 0x3ff);
    p.q1 = (int)((v >> 40) & 0xfff);
    p.q2 = (int)((v >> 52) & 0xfff);
    return p;
}

// one power-spectrum Jacobian element of a single edge record (compute type TC)
template <typename TC>
__device__ __forceinline__ TC ps_elem(const TC* rp, const RecLayout& R, int s, int d, unsigned long long plan,
                                      const TC* prefs) {
    const FPlan p = unpack_plan(plan);
    const TC rv = rp[d];
    const TC* B = rp + R.offB;
    const TC* Bt = rp + R.offBt;
    const TC* U = rp + R.offU;
    const TC* V = rp + R.offV + d * R.SNL;
    TC v = TC(0);
    if (p.j == s) v += rv * Bt[p.a1] * U[p.q1] + B[p.a1] * V[p.q1];
    if (p.jd == s) v += rv * Bt[p.a2] * U[p.q2] + B[p.a2] * V[p.q2];
    return prefs[p.l] * v;
}

// VEC: the row length is a multiple of 4 elements and the output is 16-byte aligned.
// TC: arithmetic type of the single-edge rows and the element type of the records.
template <typename TO, typename TC, bool VEC>
__global__ __launch_bounds__(256) void dense_kernel(cusoap_basis_t b, int mode, int n_atoms,
                                                    const int32_t* __restrict__ atom_sp,
                                                    const int32_t* __restrict__ center_attach,
                                                    const int64_t* __restrict__ offsets,
                                                    const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ inv_atom,
                                                    const unsigned long long* __restrict__ rowkeys,
                                                    const int32_t* __restrict__ chunk_lo, int n_inc, int chunks,
                                                    const uint32_t* __restrict__ ft, int n_feat,
                                                    const TC* __restrict__ rec, TO* __restrict__ out) {
    __shared__ unsigned long long ckeys[kChunkKeys];
    __shared__ int rstart[kRowsPerBlock + 1];
    __shared__ int nzrows[kRowsPerBlock];
    __shared__ int nnz;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = b.n_max, S = b.n_species;
    const RecLayout R = rec_layout(S, n, b.l_max, mode);
    double* prefs = (double*)smem;                                   // [16]
    TC* prefc = (TC*)(prefs + 16);                                   // [16]
    unsigned long long* plan = (unsigned long long*)(prefs + 32);    // [n_feat] (ps)
    uint32_t* fts = (uint32_t*)(plan + n_feat);                      // [n_feat] (ps)
    TC* recs = (TC*)(fts + n_feat + (n_feat & 3));                   // [4][stride]
    const int c = blockIdx.x / chunks, ch = blockIdx.x - c * chunks;
    const int r0 = ch * kRowsPerBlock;
    const int r1 = (r0 + kRowsPerBlock < n_inc) ? r0 + kRowsPerBlock : n_inc;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int lo = chunk_lo[(size_t)c * (chunks + 1) + ch], hi = chunk_lo[(size_t)c * (chunks + 1) + ch + 1];
    const bool in_lds = hi - lo <= kChunkKeys;
    const unsigned long long* keys = in_lds ? ckeys : rowkeys + e0 + lo;
    if (in_lds)
        for (int i = threadIdx.x; i < hi - lo; i += blockDim.x) ckeys[i] = rowkeys[e0 + lo + i];
    if (mode == CUSOAP_MODE_PS) {
        for (int f = threadIdx.x; f < n_feat; f += blockDim.x) {
            const uint32_t v = ft[f];
            fts[f] = v;
            const int j = (int)(v & 0xffu), jd = (int)((v >> 8) & 0xffu);
            const int nn = (int)((v >> 16) & 0x3fu), nd = (int)((v >> 22) & 0x3fu), l = (int)(v >> 28);
            plan[f] = pack_plan(j, jd, l, nn * R.L1 + l, nd * R.L1 + l, (jd * n + nd) * R.L1 + l,
                                (j * n + nn) * R.L1 + l);
        }
    }
    if ((int)threadIdx.x <= b.l_max) {
        const double pf = sqrt(8.0 * M_PI * M_PI / (2 * threadIdx.x + 1));
        prefs[threadIdx.x] = pf;
        prefc[threadIdx.x] = (TC)pf;
    }
    int att_row = -1;
    if (center_attach) {
        const int aa = center_attach[c];
        if (aa >= 0) att_row = inv_atom[aa];
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= r1 - r0; i += blockDim.x) rstart[i] = lower_row(keys, 0, hi - lo, r0 + i);
    __syncthreads();
    const int row_len = 3 * n_feat;
    // 1) every run of consecutive rows without an edge: one flat block-wide store
    if (threadIdx.x == 0) nnz = 0;
    {
        int r = r0;
        while (r < r1) {
            while (r < r1 && (rstart[r - r0] != rstart[r - r0 + 1] || r == att_row)) ++r;
            const int rs = r;
            while (r < r1 && rstart[r - r0] == rstart[r - r0 + 1] && r != att_row) ++r;
            if (r > rs) {
                TO* dst = out + ((size_t)c * n_inc + rs) * (size_t)row_len;
                const size_t cnt = (size_t)(r - rs) * row_len;
                if (VEC) {
                    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
                    const u32x4 z = {0u, 0u, 0u, 0u};
                    u32x4* p = (u32x4*)dst;
                    const size_t nv = cnt * sizeof(TO) / 16;
                    for (size_t i = threadIdx.x; i < nv; i += blockDim.x) p[i] = z;
                } else {
                    for (size_t i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = TO(0);
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)(r1 - r0)) {
        const int r = r0 + threadIdx.x;
        if (rstart[r - r0] != rstart[r - r0 + 1] || r == att_row) nzrows[atomicAdd(&nnz, 1)] = r;
    }
    __syncthreads();
    // 2) rows with edges: one wavefront per row
    TC* myrec = recs + wv * R.stride;
    for (int i = wv; i < nnz; i += 4) {
        const int r = nzrows[i];
        TO* dst = out + ((size_t)c * n_inc + r) * (size_t)row_len;
        const int q0 = rstart[r - r0];
        const int q1 = rstart[r - r0 + 1];
        if (q1 - q0 == 1 && r != att_row) {
            // one edge: stage its record in LDS (wavefront-private slot)
            const int64_t e = e0 + (int64_t)(keys[q0] & 0xffffffffull);
            const TC* rp = rec + (size_t)e * R.stride;
            for (int k = lane; k < R.stride; k += 64) myrec[k] = rp[k];
            wave_sync();
            const int s = atom_sp[nbr[e] % n_atoms];
            if (mode == CUSOAP_MODE_PS && VEC) {
                int t = 4 * lane;
                int d = t / n_feat, f = t - d * n_feat;
                for (; t < row_len; t += 256) {
                    TC v[4];
                    int dd = d, ff = f;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (ff >= n_feat) {
                            ff -= n_feat;
                            ++dd;
                        }
                        v[k] = ps_elem<TC>(myrec, R, s, dd, plan[ff], prefc);
                        ++ff;
                    }
                    store4<TO>(dst + t, (double)v[0], (double)v[1], (double)v[2], (double)v[3]);
                    f += 256;
                    while (f >= n_feat) {
                        f -= n_feat;
                        ++d;
                    }
                }
            } else if (mode == CUSOAP_MODE_PS) {
                for (int t = lane; t < row_len; t += 64) {
                    const int d = t / n_feat, f = t - d * n_feat;
                    dst[t] = (TO)ps_elem<TC>(myrec, R, s, d, plan[f], prefc);
                }
            } else {
                const TC* rpd = rec + (size_t)e * R.stride;
                for (int t = lane; t < row_len; t += 64)
                    dst[t] = (TO)elem_value(rpd, R, mode, n, s, t, n_feat, fts, prefs);
            }
            wave_sync();
            continue;
        }
        // several edges (periodic images) and/or the attach row: float64 sums of the global records
        for (int t = lane; t < row_len; t += 64) {
            double v = 0.0;
            for (int q = q0; q < q1; ++q) {
                const int64_t e = e0 + (int64_t)(keys[q] & 0xffffffffull);
                v += elem_value(rec + (size_t)e * R.stride, R, mode, n, atom_sp[nbr[e] % n_atoms], t, n_feat,
                                fts, prefs);
            }
            if (r == att_row)
                for (int64_t e = e0; e < e1; ++e)
                    v -= elem_value(rec + (size_t)e * R.stride, R, mode, n, atom_sp[nbr[e] % n_atoms], t,
                                    n_feat, fts, prefs);
            dst[t] = (TO)v;
        }
    }
}

// Power-spectrum rows of the GTO basis when every output row has exactly one edge (one
// image, no attached center: the reference's closed-form case): the wavefront that
// stores a row first forms that edge's record in LDS, as records_kernel does (float64
// arithmetic in the same order, stored as TC), so no record array goes through HBM.
// Workgroup (center, 32-row chunk), zero runs and row stores as dense_kernel (VEC form).
template <int LMAX, typename TO, typename TC>
__global__ __launch_bounds__(256) void dense_fused_kernel(cusoap_basis_t b, const double* __restrict__ pos,
                                                          int n_atoms, const int32_t* __restrict__ atom_sp,
                                                          const double* __restrict__ shifts,
                                                          const double* __restrict__ centers,
                                                          const int64_t* __restrict__ offsets,
                                                          const int32_t* __restrict__ nbr,
                                                          const double* __restrict__ coeffs,
                                                          const unsigned long long* __restrict__ rowkeys,
                                                          const int32_t* __restrict__ chunk_lo, int n_inc, int chunks,
                                                          const uint32_t* __restrict__ ft, int n_feat,
                                                          TO* __restrict__ out) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    __shared__ unsigned long long ckeys[kChunkKeys];
    __shared__ int rstart[kRowsPerBlock + 1];
    __shared__ int nzrows[kRowsPerBlock];
    __shared__ int nnz;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = b.n_max, S = b.n_species, NL = n * L1;
    const RecLayout R = rec_layout(S, n, LMAX, CUSOAP_MODE_PS);
    double* prefs = (double*)smem;                                   // [16]
    TC* prefc = (TC*)(prefs + 16);                                   // [16]
    double* Cs = prefs + 32;                                         // [S n L2] the center's coefficients
    double* Yw = Cs + S * n * L2;                                    // [4][4 L2] harmonics, gradients
    double* Qw = Yw + 16 * L2;                                       // [4][NL]   GTO primitives
    unsigned long long* plan = (unsigned long long*)(Qw + 4 * NL);   // [n_feat]
    TC* recs = (TC*)(plan + n_feat);                                 // [4][stride]
    const int c = blockIdx.x / chunks, ch = blockIdx.x - c * chunks;
    const int r0 = ch * kRowsPerBlock;
    const int r1 = (r0 + kRowsPerBlock < n_inc) ? r0 + kRowsPerBlock : n_inc;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e0 = offsets[c];
    const int lo = chunk_lo[(size_t)c * (chunks + 1) + ch], hi = chunk_lo[(size_t)c * (chunks + 1) + ch + 1];
    const bool in_lds = hi - lo <= kChunkKeys;
    const unsigned long long* keys = in_lds ? ckeys : rowkeys + e0 + lo;
    if (in_lds)
        for (int i = threadIdx.x; i < hi - lo; i += blockDim.x) ckeys[i] = rowkeys[e0 + lo + i];
    for (int f = threadIdx.x; f < n_feat; f += blockDim.x) {
        const uint32_t v = ft[f];
        const int j = (int)(v & 0xffu), jd = (int)((v >> 8) & 0xffu);
        const int nn = (int)((v >> 16) & 0x3fu), nd = (int)((v >> 22) & 0x3fu), l = (int)(v >> 28);
        plan[f] = pack_plan(j, jd, l, nn * R.L1 + l, nd * R.L1 + l, (jd * n + nd) * R.L1 + l, (j * n + nn) * R.L1 + l);
    }
    if ((int)threadIdx.x <= LMAX) {
        const double pf = sqrt(8.0 * M_PI * M_PI / (2 * threadIdx.x + 1));
        prefs[threadIdx.x] = pf;
        prefc[threadIdx.x] = (TC)pf;
    }
    for (int i = threadIdx.x; i < S * n * L2; i += blockDim.x) Cs[i] = coeffs[(size_t)c * S * n * L2 + i];
    __syncthreads();
    for (int i = threadIdx.x; i <= r1 - r0; i += blockDim.x) rstart[i] = lower_row(keys, 0, hi - lo, r0 + i);
    __syncthreads();
    const int row_len = 3 * n_feat;
    // 1) every run of consecutive rows without an edge: one flat block-wide store
    if (threadIdx.x == 0) nnz = 0;
    {
        int r = r0;
        while (r < r1) {
            while (r < r1 && rstart[r - r0] != rstart[r - r0 + 1]) ++r;
            const int rs = r;
            while (r < r1 && rstart[r - r0] == rstart[r - r0 + 1]) ++r;
            if (r > rs) {
                typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
                const u32x4 z = {0u, 0u, 0u, 0u};
                u32x4* p = (u32x4*)(out + ((size_t)c * n_inc + rs) * (size_t)row_len);
                const size_t nv = (size_t)(r - rs) * row_len * sizeof(TO) / 16;
                for (size_t i = threadIdx.x; i < nv; i += blockDim.x) p[i] = z;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)(r1 - r0)) {
        const int r = r0 + threadIdx.x;
        if (rstart[r - r0] != rstart[r - r0 + 1]) nzrows[atomicAdd(&nnz, 1)] = r;
    }
    __syncthreads();
    // 2) rows with an edge: one wavefront per row, its record formed in the wavefront's slot
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const double inv_na = 1.0 / (double)n_atoms;
    TC* myrec = recs + wv * R.stride;
    double* Y = Yw + wv * 4 * L2;
    double* Q = Qw + wv * NL;
    for (int i = wv; i < nnz; i += 4) {
        const int r = nzrows[i];
        TO* dst = out + ((size_t)c * n_inc + r) * (size_t)row_len;
        const int64_t e = e0 + (int64_t)(keys[rstart[r - r0]] & 0xffffffffull);
        int img, a;
        split_ext(nbr[e], n_atoms, inv_na, img, a);
        const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
        const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
        const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
        const double r2 = (rx * rx + ry * ry) + rz * rz;
        const double rr = sqrt(r2);
        double w = 1.0, dw = 0.0;
        if (b.weighting != CUSOAP_W_NONE) weighting(b, rr, w, dw);
        const double dwr = rr > 0.0 ? dw / rr : 0.0;
        if (lane == 0)
            solid_harmonics<LMAX, true, double>(rx, ry, rz, [&](int lm, double v, double gx, double gy, double gz) {
                Y[lm] = v;
                Y[L2 + lm] = gx;
                Y[2 * L2 + lm] = gy;
                Y[3 * L2 + lm] = gz;
            });
        for (int t = lane; t < NL; t += 64) {
            const int k = t / L1, l = t - k * L1;
            Q[t] = b.gto_pref[l * n + k] * exp(-b.gto_gamma[l * n + k] * r2);
        }
        if (lane < 3) myrec[lane] = (TC)(lane == 0 ? rx : (lane == 1 ? ry : rz));
        wave_sync();
        for (int t = lane; t < NL; t += 64) {
            const int nn = t / L1, l = t - nn * L1;
            const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
            double B = 0.0, Bh = 0.0;
            for (int k = 0; k < n; ++k) {
                const double q = Q[k * L1 + l];
                B += beta[k] * q;
                Bh += beta[k] * (-2.0 * b.gto_gamma[l * n + k]) * q;
            }
            myrec[R.offB + t] = (TC)(w * B);
            myrec[R.offBt + t] = (TC)(w * Bh + dwr * B);
        }
        for (int t = lane; t < 4 * R.SNL; t += 64) {
            const int comp = t / R.SNL, rem = t - comp * R.SNL;
            const int kq = rem / L1, l = rem - kq * L1;
            const double* yv = Y + comp * L2 + l * l;
            const double* cr = Cs + kq * L2 + l * l;
            double acc = 0.0;
            for (int m = 0; m < 2 * l + 1; ++m) acc += yv[m] * cr[m];
            myrec[R.offU + t] = (TC)acc;
        }
        wave_sync();
        const int s = atom_sp[a];
        int t = 4 * lane;
        int d = t / n_feat, f = t - d * n_feat;
        for (; t < row_len; t += 256) {
            TC v[4];
            int dd = d, ff = f;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (ff >= n_feat) {
                    ff -= n_feat;
                    ++dd;
                }
                v[k] = ps_elem<TC>(myrec, R, s, dd, plan[ff], prefc);
                ++ff;
            }
            store4<TO>(dst + t, (double)v[0], (double)v[1], (double)v[2], (double)v[3]);
            f += 256;
            while (f >= n_feat) {
                f -= n_feat;
                ++d;
            }
        }
        wave_sync();
    }
}

template <int LMAX, typename TO, typename TC>
int launch_dense_fused(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                       const double* shifts, const double* centers, const int64_t* offsets, const int32_t* nbr,
                       const double* coeffs, const unsigned long long* rowkeys, const int32_t* chunk_lo, int n_centers,
                       int n_inc, const uint32_t* ft, int n_feat, void* out, hipStream_t st) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const RecLayout R = rec_layout(b.n_species, b.n_max, LMAX, CUSOAP_MODE_PS);
    if (R.SNL >= 4096 || R.NL >= 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: too many radial channels");
    const size_t smem = sizeof(double) * (32 + (size_t)b.n_species * b.n_max * L2 + 16 * L2 + 4 * (size_t)R.NL) +
                        sizeof(unsigned long long) * (size_t)n_feat + sizeof(TC) * 4 * (size_t)R.stride;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: record too large for LDS");
    const int chunks = (n_inc + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t grid = (int64_t)n_centers * chunks;
    if (grid >= ((int64_t)1 << 31)) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: output too large");
    auto kern = dense_fused_kernel<LMAX, TO, TC>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("jacobian_dense", st);
    kern<<<(unsigned)grid, 256, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, offsets, nbr, coeffs, rowkeys,
                                            chunk_lo, n_inc, chunks, ft, n_feat, (TO*)out);
    prof_end("jacobian_dense", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

int records_tile(const RecLayout& R, int S, int n, size_t* smem) {
    const size_t fixed = (size_t)S * n * R.L2;
    const size_t per = 4 * (size_t)R.L2 + 2 * (size_t)R.NL + 8;
    size_t room = 8192 > fixed ? 8192 - fixed : 0;     // doubles in 64 KiB
    size_t te = room / per;
    // at most 32 edges per tile: ~35 KiB of LDS at n 7, l 3 leaves room for 4 workgroups
    // per CU (64-edge tiles, 2 per CU, were 25% slower)
    if (te > 32) te = 32;
    if (te < 8) te = 8;
    *smem = sizeof(double) * (fixed + te * per);
    return (int)te;
}

template <int LMAX, typename TR>
int launch_records(const cusoap_basis_t& b, int mode, const double* pos, int n_atoms, const double* shifts,
                   const double* centers, int n_centers, const int64_t* offsets, const int32_t* nbr,
                   int64_t n_edges, const double* coeffs, TR* rec, hipStream_t st) {
    const RecLayout R = rec_layout(b.n_species, b.n_max, LMAX, mode);
    size_t smem = 0;
    const int TE = records_tile(R, b.n_species, b.n_max, &smem);
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: coefficient block too large for LDS");
    // several workgroups per center (one per tile of a mean-sized center): one workgroup
    // per center left most of the GPU idle on a few thousand centers
    const int64_t mean_tiles = n_centers > 0 ? (n_edges / n_centers + TE - 1) / TE : 1;
    const dim3 grid((unsigned)n_centers, (unsigned)std::max<int64_t>(1, std::min<int64_t>(16, mean_tiles)));
    prof_begin("jacobian_edges", st);
    if (b.rbf == CUSOAP_RBF_GTO) {
        auto kern = records_kernel<LMAX, CUSOAP_RBF_GTO, TR>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        kern<<<grid, 256, smem, st>>>(b, mode, TE, pos, n_atoms, shifts, centers, offsets, nbr, coeffs, rec);
    } else {
        auto kern = records_kernel<LMAX, CUSOAP_RBF_POLY, TR>;
        if (smem > 64 * 1024)
            CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        kern<<<grid, 256, smem, st>>>(b, mode, TE, pos, n_atoms, shifts, centers, offsets, nbr, coeffs, rec);
    }
    prof_end("jacobian_edges", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <typename TO, typename TC, bool VEC>
int launch_dense(const cusoap_basis_t& b, int mode, int n_atoms, const int32_t* atom_sp, const int32_t* center_attach,
                 const int64_t* offsets, const int32_t* nbr, const int32_t* inv_atom,
                 const unsigned long long* rowkeys, const int32_t* chunk_lo, int n_centers, int n_inc,
                 const uint32_t* ft, int n_feat, const TC* rec, void* out, hipStream_t st) {
    const RecLayout R = rec_layout(b.n_species, b.n_max, b.l_max, mode);
    if (R.SNL >= 4096 || R.NL >= 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: too many radial channels");
    const size_t smem = sizeof(double) * 32 + (sizeof(unsigned long long) + sizeof(uint32_t)) * (size_t)n_feat +
                        sizeof(uint32_t) * 4 + sizeof(TC) * 4 * (size_t)R.stride;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: record too large for LDS");
    const int chunks = (n_inc + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t grid = (int64_t)n_centers * chunks;
    if (grid >= ((int64_t)1 << 31)) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: output too large");
    auto kern = dense_kernel<TO, TC, VEC>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("jacobian_dense", st);
    kern<<<(unsigned)grid, 256, smem, st>>>(b, mode, n_atoms, atom_sp, center_attach, offsets, nbr, inv_atom, rowkeys,
                                            chunk_lo, n_inc, chunks, ft, n_feat, rec, (TO*)out);
    prof_end("jacobian_dense", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

struct JacWs {
    double* rec;
    unsigned long long* rowkeys;
    int32_t* kcount;
    int32_t* chunk_lo;
    int32_t* err;
    size_t bytes;
};

JacWs carve_jac(void* base, const RecLayout& R, int64_t n_edges, int n_centers, int n_inc) {
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
    const size_t chunks = (size_t)(std::max(n_inc, 1) + kRowsPerBlock - 1) / kRowsPerBlock;
    w.chunk_lo = (int32_t*)take(sizeof(int32_t) * (size_t)std::max(n_centers, 1) * (chunks + 1));
    w.err = (int32_t*)take(sizeof(int32_t) * 4);
    w.bytes = off;
    return w;
}

}  // namespace

size_t jacobian_ws_bytes(const cusoap_basis_t* b, int mode, int64_t n_edges, int n_centers, int n_inc) {
    const RecLayout R = rec_layout(b->n_species, b->n_max, b->l_max, mode);
    return carve_jac(nullptr, R, n_edges, n_centers, n_inc).bytes;
}

int jacobian_dense_impl(const cusoap_basis_t* b, int mode, const double* pos, int n_atoms, const int32_t* atom_sp,
                        const double* shifts, int n_img, const double* centers, int n_centers,
                        const int32_t* center_attach, const int64_t* offsets, const int32_t* nbr, int64_t n_edges,
                        const double* coeffs, const int32_t* inv_atom, int n_inc, const uint32_t* ft, int n_feat,
                        void* ws, size_t ws_bytes, void* out, int odt, int cdt, hipStream_t st) {
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: n_max must be 1..16 and l_max 0..9");
    if (mode != CUSOAP_MODE_PS && mode != CUSOAP_MODE_CC) return fail(CUSOAP_ERR_ARG, "jacobian: unknown mode");
    if (b->n_species > 255 || b->n_max > 63) return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: feature table range");
    const RecLayout R = rec_layout(b->n_species, b->n_max, b->l_max, mode);
    JacWs w = carve_jac(ws, R, n_edges, n_centers, n_inc);
    if (ws_bytes < w.bytes) return fail(CUSOAP_ERR_WORKSPACE, "jacobian: workspace too small");
    if (n_centers == 0 || n_inc == 0 || n_feat == 0) return CUSOAP_OK;
    CUSOAP_HIP(hipMemsetAsync(w.err, 0, sizeof(int32_t), st));
    const int chunks = (n_inc + kRowsPerBlock - 1) / kRowsPerBlock;
    rowkey_kernel<<<n_centers, 256, 0, st>>>(n_atoms, offsets, nbr, inv_atom, w.rowkeys, w.kcount, chunks,
                                             w.chunk_lo, w.err);
    CUSOAP_LAUNCH_CHECK();
    int rc = CUSOAP_OK;
    // records in float32 when the rows are computed in float32 (half the record traffic)
    const bool f32 = odt == CUSOAP_DT_F32 && cdt == CUSOAP_DT_F32;
    // 4-element groups per lane: every row length must be a multiple of 4 elements
    const bool vec = ((3 * (size_t)n_feat) % 4 == 0) && (((uintptr_t)out & 15) == 0);
    if (n_img == 1 && !center_attach && b->rbf == CUSOAP_RBF_GTO && mode == CUSOAP_MODE_PS && vec) {
        // one edge per output row: each row's record is formed where the row is stored
#define CUSOAP_FUSED(L, TO, TC)                                                                                   \
    launch_dense_fused<L, TO, TC>(*b, pos, n_atoms, atom_sp, shifts, centers, offsets, nbr, coeffs, w.rowkeys,      \
                                  w.chunk_lo, n_centers, n_inc, ft, n_feat, out, st)
#define CUSOAP_CASE(L)                                                                                           \
    case L:                                                                                                      \
        rc = f32 ? CUSOAP_FUSED(L, float, float)                                                                 \
                 : (odt == CUSOAP_DT_F32 ? CUSOAP_FUSED(L, float, double) : CUSOAP_FUSED(L, double, double));    \
        break;
        switch (b->l_max) {
            CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
            CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
            default: return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: l_max out of range");
        }
#undef CUSOAP_CASE
#undef CUSOAP_FUSED
        if (rc) return rc;
        int32_t herr = 0;
        CUSOAP_HIP(hipMemcpyAsync(&herr, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        CUSOAP_HIP(hipStreamSynchronize(st));
        if (herr) return fail(CUSOAP_ERR_CAPACITY, "jacobian: a center has more than 4096 included edges");
        return CUSOAP_OK;
    }
#define CUSOAP_CASE(L) \
    case L: \
        rc = f32 ? launch_records<L, float>(*b, mode, pos, n_atoms, shifts, centers, n_centers, offsets, nbr, n_edges, \
                                            coeffs, (float*)w.rec, st) \
                 : launch_records<L, double>(*b, mode, pos, n_atoms, shifts, centers, n_centers, offsets, nbr, n_edges, \
                                             coeffs, w.rec, st); \
        break;
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: return fail(CUSOAP_ERR_UNSUPPORTED, "jacobian: l_max out of range");
    }
#undef CUSOAP_CASE
    if (rc) return rc;
#define CUSOAP_DENSE(TO, TC, V) \
    launch_dense<TO, TC, V>(*b, mode, n_atoms, atom_sp, center_attach, offsets, nbr, inv_atom, w.rowkeys, w.chunk_lo, \
                            n_centers, n_inc, ft, n_feat, (const TC*)w.rec, out, st)
    if (f32) rc = vec ? CUSOAP_DENSE(float, float, true) : CUSOAP_DENSE(float, float, false);
    else if (odt == CUSOAP_DT_F32) rc = vec ? CUSOAP_DENSE(float, double, true) : CUSOAP_DENSE(float, double, false);
    else rc = vec ? CUSOAP_DENSE(double, double, true) : CUSOAP_DENSE(double, double, false);
#undef CUSOAP_DENSE
    if (rc) return rc;
    int32_t herr = 0;
    CUSOAP_HIP(hipMemcpyAsync(&herr, w.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CUSOAP_HIP(hipStreamSynchronize(st));
    if (herr) return fail(CUSOAP_ERR_CAPACITY, "jacobian: a center has more than 4096 included edges");
    return CUSOAP_OK;
}

}  // namespace cusoap
