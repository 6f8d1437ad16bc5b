// Chebyshev radial form on the matrix cores (basis.rad_cheb): density coefficients.
//
// No reference counterpart for the arithmetic; the result is the coefficient array
// of _coefficients_gto / _add_self_terms (soap_generator.py:1294-1391, 821-850) for
// the radial functions that basis.rad_cheb represents.  Each radial function that
// multiplies the unit-vector harmonics, f_nl(r) = r^l B_nl(r), is a Chebyshev series
// sum_k a_lnk T_k(t), t = 2 r / cutoff - 1 (cusoap_amd/chebyshev.py).  Then
//
//   c[s][n][lm] = sum_{e of species s} f_nl(r_e) Y_lm(u_e) = sum_k a_lnk Q_s[k][lm],
//   Q_s[k][lm]  = sum_{e of species s} T_k(t_e) Y_lm(u_e),
//
// and Q_s is a dense product over the edges with no l-dependence left in it: rows k,
// columns lm, contraction over edges.  It runs as v_mfma_f32_16x16x32_f16 products.
//
// One wavefront (workgroup) per center, tiles of 64 edges, lane = edge:
//   1. geometry in float64, t and the unit vector in float32, T_k by the three-term
//      recurrence (fma), the real harmonics Y_lm(u) (solid_harmonics at |u| = 1);
//   2. both operands scaled by 2^14 and split into fp16 hi + lo (lo = fp16(v - hi)), so
//      that hi*hi + hi*lo + lo*hi carries ~22 bits, and every value stays in the
//      normal fp16 range or below 2^-14 absolute (2^-28 of the scaled maximum);
//      stored to LDS rows [k][position] and [lm][position], 16-byte chunks of a row
//      XOR-swizzled by the row, positions compacted by species (ballot + mbcnt);
//   3. per 32 positions: A = T rows (k = 16 kt + lane % 16, positions 8 (lane / 16)..+7),
//      B = Y rows (lm = 16 lt + lane % 16), three products per (kt, lt) tile into the
//      float32 accumulators of each species present in the block; positions of other
//      species are masked out of A, positions past the tile's edges hold zeros;
//   4. after the last tile the accumulator lane (lm, k = 16 kt + 4 (lane / 16) + i)
//      forms its partial sum_k a_lnk Q[k][lm] for 8 n, the four lane groups are added
//      (xor 16, xor 32, fixed order) and the self term is added.
// Operand layouts of v_mfma_f32_16x16x32_f16: A[i][k] in lane i + 16 (k / 8), element
// k % 8; B[k][j] in lane j + 16 (k / 8), element k % 8; D[i][j] in lane j + 16 (i / 4),
// element i % 4.  The contraction index only has to agree between A and B.
#include "common.hpp"

namespace cusoap {

namespace {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int kCE = 64;                                   // edges per tile, lane = edge
constexpr int kCT = 32;                                   // Chebyshev terms per interval
constexpr float kOpScale = 16384.0f;                      // operands in [-2^14, 2^14]
constexpr float kOpUnscale = 1.0f / (16384.0f * 16384.0f);

__host__ __device__ constexpr int lm_tiles(int lmax) { return ((lmax + 1) * (lmax + 1) + 15) / 16; }

// element offset of (row, position) in an LDS operand block of 64-position rows
// (128 bytes); the eight 16-byte chunks of a row are XOR-swizzled by the row index, so
// the 16 rows of one operand read fall on different banks
__device__ __forceinline__ int opnd(int row, int p) { return row * kCE + ((((p >> 3) ^ row) & 7) << 3) + (p & 7); }

__device__ __forceinline__ void put_hilo(_Float16* hi, _Float16* lo, int off, float v) {
    const _Float16 h = (_Float16)v;
    hi[off] = h;
    lo[off] = (_Float16)(v - (float)h);
}

// 0xFFFF in half j of the four dwords where bit j of `bits` is set
__device__ __forceinline__ uint4 half_mask(uint32_t bits) {
    uint32_t m[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        m[q] = (((bits >> (2 * q)) & 1u) ? 0x0000FFFFu : 0u) | (((bits >> (2 * q + 1)) & 1u) ? 0xFFFF0000u : 0u);
    return make_uint4(m[0], m[1], m[2], m[3]);
}

__device__ __forceinline__ int degree_of(int lm, int lmax) {
    int l = (int)sqrtf((float)lm + 0.5f);
    return l < lmax ? l : lmax;
}

// T_{8q} .. T_{8q+7} of t (q = 0..3): the start pair (T_{8q}, T_{8q+1}) from the product
// rule T_{m+n} = 2 T_m T_n - T_{|m-n|}, then the three-term recurrence (fma)
__device__ __forceinline__ void cheb_rows8(float t, int q, float (&T)[8]) {
    const float t2 = fmaf(2.0f * t, t, -1.0f);
    float a = 1.0f, b = t;                                    // T_{8q}, T_{8q+1}
    if (q > 0) {
        const float t3 = fmaf(2.0f * t, t2, -t);
        const float t4 = fmaf(2.0f * t2, t2, -1.0f);
        const float t5 = fmaf(2.0f * t2, t3, -t);
        const float t8 = fmaf(2.0f * t4, t4, -1.0f);
        const float t9 = fmaf(2.0f * t4, t5, -t);
        a = t8;
        b = t9;
        if (q > 1) {
            const float t16 = fmaf(2.0f * t8, t8, -1.0f);
            const float t17 = fmaf(2.0f * t8, t9, -t);
            a = t16;
            b = t17;
            if (q > 2) {
                const float t7 = fmaf(2.0f * t4, t3, -t);
                a = fmaf(2.0f * t16, t8, -t8);                // T24
                b = fmaf(2.0f * t16, t9, -t7);                // T25
            }
        }
    }
    T[0] = a;
    T[1] = b;
#pragma unroll
    for (int k = 2; k < 8; ++k) T[k] = fmaf(2.0f * t, T[k - 1], -T[k - 2]);
}

template <int LMAX, int SB>
__global__ __launch_bounds__(256) void cheb_coeff_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                         const int32_t* __restrict__ atom_sp,
                                                         const double* __restrict__ shifts,
                                                         const double* __restrict__ centers,
                                                         const int32_t* __restrict__ self_sp,
                                                         const int64_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ nbr, int create_path, int n_centers,
                                                         int per_xcd, float* __restrict__ out) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1, LT = lm_tiles(LMAX);
    constexpr int NG = 2 * SB;                                // (species, interval) groups
    __shared__ __align__(16) _Float16 Th[kCT * kCE];
    __shared__ __align__(16) _Float16 Tl[kCT * kCE];
    __shared__ __align__(16) _Float16 Yh[(L2 + 1) * kCE];    // row L2: zeros (columns lm >= L2)
    __shared__ __align__(16) _Float16 Yl[(L2 + 1) * kCE];
    __shared__ __align__(16) float4 geo[4][kCE];              // (t, unit vector) by tile and position
    __shared__ int gcnt[4][NG + 1];                           // edges per group, then the tile total
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, cr = lane & 15;
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront
    // consecutive workgroups go to different XCDs: give each XCD a contiguous range of
    // centers, whose neighbourhoods overlap in its L2
    const int c = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (c >= n_centers) return;
    const int S = b.n_species, n = b.n_max;
    if (tid < kCE) {
        Yh[L2 * kCE + tid] = (_Float16)0.0f;
        Yl[L2 * kCE + tid] = (_Float16)0.0f;
    }
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int self_s = (create_path && self_sp) ? self_sp[c] : -1;
    const int cslot = self_sp[c] >= 0 ? self_sp[c] : S;
    const double inv_na = 1.0 / (double)n_atoms;
    const float rs = (float)b.cheb_split, tsc0 = 2.0f / rs, tsc1 = (float)(2.0 / (b.cutoff - b.cheb_split));
    // wavefront q owns columns lm = 16 q .. 16 q + 15 of the product (waves q >= LT
    // only build operands)
    const bool mine = q < LT;
    const int lt = q < LT ? q : 0;

    f4_t acc[NG][2];
#pragma unroll
    for (int gi = 0; gi < NG; ++gi)
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) acc[gi][kt] = f4_t{0.0f, 0.0f, 0.0f, 0.0f};

    for (int64_t r0 = e0; r0 < e1; r0 += 4 * kCE) {
        // ---- geometry of a round of four tiles: wavefront q, tile q, lane = edge; the
        // edge goes to position p of its tile, compacted by (species, interval)
        {
            const int64_t e = r0 + q * kCE + lane;
            int grp = -1;
            float t = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
            if (e < e1) {
                int img, a;
                split_ext(nbr[e], n_atoms, inv_na, img, a);
                const float rx = (float)((pos[3 * (size_t)a] + shifts[3 * img]) - cx);
                const float ry = (float)((pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy);
                const float rz = (float)((pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz);
                const float r2 = (rx * rx + ry * ry) + rz * rz;
                if (!(create_path && !(r2 > 1e-24f))) {
                    const float ir = r2 > 1e-24f ? __frsqrt_rn(r2) : 0.0f;
                    const float r = r2 * ir;
                    const int outer = r >= rs;
                    grp = 2 * atom_sp[a] + outer;
                    t = outer ? fmaf(r - rs, tsc1, -1.0f) : fmaf(r, tsc0, -1.0f);
                    ux = rx * ir;
                    uy = ry * ir;
                    uz = rz * ir;
                }
            }
            int p = -1, nv = 0;
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                const uint64_t m = __ballot(grp == gi);
                const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (grp == gi) p = nv + rank;
                if (lane == 0) gcnt[q][gi] = __popcll(m);
                nv += __popcll(m);
            }
            if (lane == 0) gcnt[q][NG] = nv;
            if (p >= 0) geo[q][p] = make_float4(t, ux, uy, uz);
        }
        __syncthreads();
        const int ntile = (int)((e1 - r0 + kCE - 1) / kCE) < 4 ? (int)((e1 - r0 + kCE - 1) / kCE) : 4;
        for (int j = 0; j < ntile; ++j) {
            // ---- operands of tile j, lane = position: wavefront q writes the Chebyshev rows
            // 8q .. 8q+7 and the harmonics of orders m = q (mod 4)
            const int nv = gcnt[j][NG];
            {
                const float4 gv = geo[j][lane];
                const bool ok = lane < nv;
                const float sc = ok ? kOpScale : 0.0f;
                const float t = ok ? gv.x : 0.0f;
                float T[8];
                cheb_rows8(t, q, T);
#pragma unroll
                for (int k = 0; k < 8; ++k) put_hilo(Th, Tl, opnd(8 * q + k, lane), T[k] * sc);
                solid_harmonics_stripe<LMAX, float>(ok ? gv.y : 0.0f, ok ? gv.z : 0.0f, ok ? gv.w : 0.0f, q,
                                                    [&](int l, int m, float v) {
                                                        put_hilo(Yh, Yl, opnd(l * l + l + m, lane), v * sc);
                                                    });
            }
            __syncthreads();
            // ---- products: wavefront q, columns lm = 16 q + lane % 16
            if (mine) {
                int beg[NG], cnt[NG];
                {
                    int acc_n = 0;
#pragma unroll
                    for (int gi = 0; gi < NG; ++gi) {
                        beg[gi] = acc_n;
                        cnt[gi] = gcnt[j][gi];
                        acc_n += cnt[gi];
                    }
                }
                const int nb = (nv + 31) >> 5;
                for (int blk = 0; blk < nb; ++blk) {
                    const int ch = 4 * blk + g;
                    uint4 ah[2], al[2];
                    h8_t bh, bl;
#pragma unroll
                    for (int kt = 0; kt < 2; ++kt) {
                        const int row = 16 * kt + cr;
                        const int off = row * kCE + (((ch ^ row) & 7) << 3);
                        ah[kt] = *(const uint4*)(Th + off);
                        al[kt] = *(const uint4*)(Tl + off);
                    }
                    {
                        int row = 16 * lt + cr;
                        row = row < L2 ? row : L2;
                        const int off = row * kCE + (((ch ^ row) & 7) << 3);
                        bh = *(const h8_t*)(Yh + off);
                        bl = *(const h8_t*)(Yl + off);
                    }
                    const int b0 = 32 * blk, p0 = b0 + 8 * g;
#pragma unroll
                    for (int gi = 0; gi < NG; ++gi) {
                        const int lo = beg[gi], hi = beg[gi] + cnt[gi];
                        if (cnt[gi] == 0 || hi <= b0 || lo >= b0 + 32) continue;
                        uint4 mh[2], ml[2];
                        if (lo <= b0 && hi >= b0 + 32) {
#pragma unroll
                            for (int kt = 0; kt < 2; ++kt) {
                                mh[kt] = ah[kt];
                                ml[kt] = al[kt];
                            }
                        } else {
                            const int a0 = min(max(lo - p0, 0), 8), a1 = min(max(hi - p0, 0), 8);
                            const uint32_t bits = ((1u << a1) - 1u) & ~((1u << a0) - 1u);
                            const uint4 mk = half_mask(bits);
#pragma unroll
                            for (int kt = 0; kt < 2; ++kt) {
                                mh[kt] = make_uint4(ah[kt].x & mk.x, ah[kt].y & mk.y, ah[kt].z & mk.z, ah[kt].w & mk.w);
                                ml[kt] = make_uint4(al[kt].x & mk.x, al[kt].y & mk.y, al[kt].z & mk.z, al[kt].w & mk.w);
                            }
                        }
#pragma unroll
                        for (int kt = 0; kt < 2; ++kt) {
                            const h8_t xh = __builtin_bit_cast(h8_t, mh[kt]), xl = __builtin_bit_cast(h8_t, ml[kt]);
                            acc[gi][kt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh, acc[gi][kt], 0, 0, 0);
                            acc[gi][kt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl, acc[gi][kt], 0, 0, 0);
                            acc[gi][kt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh, acc[gi][kt], 0, 0, 0);
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (!mine) return;

    // c[s][n][lm] = 2^-28 sum_{p, j} a_lnpj Q_sp[j][lm]; this lane holds j = 16 kt + 4 g + i
    const float* tab = b.rad_cheb;
    const int lm = 16 * lt + cr;
    const int dl = degree_of(lm, LMAX);
#pragma unroll
    for (int s = 0; s < SB; ++s) {
        if (s >= S) break;
        const float* Al = tab + ((size_t)(cslot * S + s) * 2 * L1 + dl) * 2 * kCT * 8;
        float cn[8];
#pragma unroll
        for (int nn = 0; nn < 8; ++nn) cn[nn] = 0.0f;
#pragma unroll
        for (int pp = 0; pp < 2; ++pp)
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int k = pp * kCT + 16 * kt + 4 * g + i;
                    const float qv = acc[2 * s + pp][kt][i];
                    const float4 a0 = *(const float4*)(Al + k * 8);
                    const float4 a1 = *(const float4*)(Al + k * 8 + 4);
                    cn[0] = fmaf(a0.x, qv, cn[0]);
                    cn[1] = fmaf(a0.y, qv, cn[1]);
                    cn[2] = fmaf(a0.z, qv, cn[2]);
                    cn[3] = fmaf(a0.w, qv, cn[3]);
                    cn[4] = fmaf(a1.x, qv, cn[4]);
                    cn[5] = fmaf(a1.y, qv, cn[5]);
                    cn[6] = fmaf(a1.z, qv, cn[6]);
                    cn[7] = fmaf(a1.w, qv, cn[7]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
        for (int nn = 0; nn < 8; ++nn) {
            cn[nn] += __shfl_xor(cn[nn], 16, 64);
            cn[nn] += __shfl_xor(cn[nn], 32, 64);
        }
        if (lm < L2) {
#pragma unroll
            for (int nn = 0; nn < 8; ++nn) {
                if ((nn >> 1) != g || nn >= n) continue;
                float v = cn[nn] * kOpUnscale;
                if (s == self_s && lm == 0) v += (float)b.self_l0[nn];
                out[(((size_t)c * S + s) * n + nn) * L2 + lm] = v;
            }
        }
    }
}

// accumulator tiles of one instantiation: kept to what fits the register file
template <int LMAX, int SB>
constexpr bool cheb_fits() { return LMAX <= 6; }

template <int LMAX, int SB>
int launch_cheb_coeff(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                      const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                      const int64_t* offsets, const int32_t* nbr, int create_path, float* out, hipStream_t st) {
    if constexpr (!cheb_fits<LMAX, SB>()) {
        return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: Chebyshev form not compiled for this l_max / species");
    } else {
        const int per = (n_centers + 7) / 8;
        prof_begin("coefficients", st);
        cheb_coeff_kernel<LMAX, SB><<<8 * per, 256, 0, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, self_sp,
                                                             offsets, nbr, create_path, n_centers, per, out);
        prof_end("coefficients", st);
        CUSOAP_LAUNCH_CHECK();
        return CUSOAP_OK;
    }
}

template <int LMAX>
int dispatch_cheb_coeff(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                        const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                        const int64_t* offsets, const int32_t* nbr, int create_path, float* out, hipStream_t st) {
#define CUSOAP_CHEB_ARGS b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st
    if (b.n_species <= 2) return launch_cheb_coeff<LMAX, 2>(CUSOAP_CHEB_ARGS);
    return launch_cheb_coeff<LMAX, 4>(CUSOAP_CHEB_ARGS);
#undef CUSOAP_CHEB_ARGS
}

}  // namespace

#define CUSOAP_CHEB_PARAMS const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp, \
    const double* shifts, const double* centers, int n_centers, const int32_t* self_sp, const int64_t* offsets, \
    const int32_t* nbr, int create_path, float* out, hipStream_t st
#define CUSOAP_CHEB_DECL(L) int cheb_coefficients_l##L(CUSOAP_CHEB_PARAMS);
CUSOAP_CHEB_DECL(0) CUSOAP_CHEB_DECL(1) CUSOAP_CHEB_DECL(2) CUSOAP_CHEB_DECL(3) CUSOAP_CHEB_DECL(4)
CUSOAP_CHEB_DECL(5) CUSOAP_CHEB_DECL(6)
#undef CUSOAP_CHEB_DECL

#ifdef CUSOAP_ONLY_L
#if CUSOAP_ONLY_L <= 6
#define CUSOAP_CHEB_DEF(L)                                                                                       \
    int cheb_coefficients_l##L(CUSOAP_CHEB_PARAMS) {                                                            \
        return dispatch_cheb_coeff<L>(b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, \
                                      create_path, out, st);                                                     \
    }
#define CUSOAP_CHEB_DEF2(L) CUSOAP_CHEB_DEF(L)
CUSOAP_CHEB_DEF2(CUSOAP_ONLY_L)
#endif
#else
// Coefficients in the Chebyshev radial form (called by coefficients_impl when
// basis.rad_cheb is set; float32 output).
int cheb_coefficients_impl(const cusoap_basis_t* b, const double* pos, int n_atoms, const int32_t* atom_sp,
                           const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                           const int64_t* offsets, const int32_t* nbr, int create_path, float* out, hipStream_t st) {
    if (!self_sp) return fail(CUSOAP_ERR_ARG, "coefficients: the Chebyshev form needs center_self_species");
    if (b->cheb_terms != kCT || !(b->cheb_split > 0.0) || !(b->cheb_split < b->cutoff))
        return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: the Chebyshev form needs cheb_terms 32 and 0 < cheb_split < cutoff");
    if (b->n_max > 8 || b->n_species > 4 || b->l_max > 6 || b->rbf != CUSOAP_RBF_GTO || b->weighting != CUSOAP_W_NONE)
        return fail(CUSOAP_ERR_UNSUPPORTED,
                    "coefficients: the Chebyshev form covers unweighted GTO bases with n_max <= 8, l_max <= 6, "
                    "at most 4 species");
    if (n_centers == 0) return CUSOAP_OK;
#define CUSOAP_CASE(L) \
    case L: return cheb_coefficients_l##L(*b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st);
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4) CUSOAP_CASE(5) CUSOAP_CASE(6)
        default: break;
    }
#undef CUSOAP_CASE
    return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: l_max out of range");
}
#endif  // CUSOAP_ONLY_L

}  // namespace cusoap
