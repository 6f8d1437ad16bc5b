// Chebyshev radial form on the matrix cores (basis.rad_cheb): density coefficients.
//
// No reference counterpart for the arithmetic; the result is the coefficient array
// of _coefficients_gto / _add_self_terms (soap_generator.py:1294-1391, 821-850) for
// the radial functions that basis.rad_cheb represents.  Each radial function that
// multiplies the unit-vector harmonics, f_nl(r) = r^l B_nl(r), is a piecewise Chebyshev
// series on two intervals p, [0, r_s] and [r_s, cutoff] (r_s = basis.cheb_split), 32 terms
// each: f_nl(r) = sum_k a_lnpk T_k(t_p(r)), t_p the interval's coordinate in [-1, 1]
// (cusoap_amd/chebyshev.py).  Then
//
//   c[s][n][lm] = sum_{e of species s} f_nl(r_e) Y_lm(u_e) = sum_{p,k} a_lnpk Q_sp[k][lm],
//   Q_sp[k][lm] = sum_{e of species s in interval p} T_k(t_p(r_e)) Y_lm(u_e),
//
// and Q_sp is a dense product over the edges with no l-dependence left in it: rows k,
// columns lm, contraction over edges.  It runs as v_mfma_f32_16x16x32_f16 products.
//
// Four wavefronts (one workgroup) per center; consecutive workgroups go to different XCDs,
// each XCD taking a contiguous range of centers.  Rounds of up to 256 edges:
//   1. geometry (wavefront q: edges 64 q + lane) in float64, t and the unit vector in
//      float32; the round's edges sorted by (species, interval) group (ballot + mbcnt),
//      each group starting on a multiple of 8 positions;
//   2. operands of 64-position tiles, double-buffered in LDS: wavefront q writes the
//      Chebyshev rows 8 q .. 8 q + 7 (product-rule starts, then the recurrence) and the
//      harmonics of the orders m in its stripe (StripeOf); both operands scaled by 2^14 and
//      split into fp16 hi + lo (lo = fp16(v - hi)), so that hi*hi + hi*lo + lo*hi carries
//      ~22 bits; rows [k][position] and [lm][position], 16-byte chunks XOR-swizzled by row;
//   3. wavefront q owns lm tile q (LmPack: whole degrees packed into 16 columns) and forms
//      Q for its columns, three products per (row tile, 32 positions) into the float32
//      accumulators of each group present; a lane's 8 positions belong to one group, so
//      positions of other groups are masked with one select; the products of tile j
//      overlap the construction of tile j + 1;
//   4. after the last round, c = sum_{p,k} a Q on the matrix cores: A = the precomputed
//      fp16 hi/lo coefficient rows (n) of each degree of the tile (basis.cheb_frag,
//      scaled per row, basis.cheb_rowinv undoing it), B = Q straight from the
//      accumulators (their lanes are the B layout for k = 4 g + (j % 4) + 16 (j / 4)),
//      columns of other degrees zeroed; then the self term.
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

// Two values as fp16 hi + lo pairs (hi rounded to nearest, lo = fp16(v - hi)): ~22 bits each.
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void put_hilo2(_Float16* hi, _Float16* lo, int off0, int off1, float v0, float v1) {
    const f2_t v = {v0, v1};
    const h2_t h = __builtin_convertvector(v, h2_t);           // one packed conversion for both
    const h2_t l = __builtin_convertvector(v - __builtin_convertvector(h, f2_t), h2_t);
    hi[off0] = h.x;
    hi[off1] = h.y;
    lo[off0] = l.x;
    lo[off1] = l.y;
}

__device__ __forceinline__ void put_hilo(_Float16* hi, _Float16* lo, int off, float v) {
    const _Float16 h = (_Float16)v;
    hi[off] = h;
    lo[off] = (_Float16)(v - (float)h);
}

// Real harmonics Y_lm(u) of the orders m of stripe `phase`, |u| = 1, the order pair (l, +m),
// (l, -m) delivered together: put1(l, v) for m = 0, put2(l, m, v_+m, v_-m) for m > 0.  The
// recurrence of solid_harmonics_stripe (associated Legendre in (z, r^2) with (x + i y)^m).
// orders m split into four stripes of near-equal value counts (l_max + 1 values for m = 0,
// 2 (l_max - m + 1) for m > 0), largest first into the lightest stripe: l_max 6 gives
// {1}, {2, 6}, {3, 5}, {0, 4}
template <int LMAX>
struct StripeOf {
    int s[LMAX + 1] = {};
    constexpr StripeOf() {
        int load[4] = {0, 0, 0, 0};
        bool done[LMAX + 1] = {};
        for (int it = 0; it <= LMAX; ++it) {
            int best = -1, cnt = -1;
            for (int m = 0; m <= LMAX; ++m) {
                const int c = m == 0 ? LMAX + 1 : 2 * (LMAX - m + 1);
                if (!done[m] && c > cnt) {
                    cnt = c;
                    best = m;
                }
            }
            int t = 0;
            for (int k = 1; k < 4; ++k)
                if (load[k] < load[t]) t = k;
            s[best] = t;
            load[t] += cnt;
            done[best] = true;
        }
    }
};

// As solid_harmonics_stripe, scaled by sc (folded into the start of the linear recurrence).
template <int LMAX, typename P1, typename P2>
__device__ __forceinline__ void harmonics_stripe_pairs(float x, float y, float z, int phase, float sc, P1&& put1,
                                                       P2&& put2) {
    constexpr StripeOf<LMAX> kStripe{};
    const float r2 = (x * x + y * y) + z * z;
    float cm = 1.0f, sm = 0.0f;
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > 0) {
            const float nc = x * cm - y * sm;
            const float ns = x * sm + y * cm;
            cm = nc;
            sm = ns;
        }
        if (kStripe.s[m] != phase) continue;
        float q2 = 0.0f, q1 = (float)kDoubleFact[m] * sc;
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            float qv;
            if (l == m) {
                qv = q1;
            } else if (l == m + 1) {
                qv = (float)(2 * m + 1) * z * q1;
            } else {
                qv = ((float)(2 * l - 1) * z * q1 - (float)(l + m - 1) * r2 * q2) * (float)(1.0 / (l - m));
            }
            if (l > m) {
                q2 = q1;
                q1 = qv;
            }
            const float K = (float)kSHNorm[l][m];
            if (m == 0)
                put1(l, K * qv);
            else
                put2(l, m, K * qv * cm, K * qv * sm);
        }
    }
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

// Columns lm of the product packed into tiles of 16 by whole degrees (first fit, degrees
// in decreasing order): l_max 6 gives {6, 1}, {5, 2}, {4, 3}, {0}, so a wavefront's
// contraction touches at most two degrees.  row[t][i] = lm of column i of tile t, or -1.
template <int LMAX>
struct LmPack {
    int ntile = 0;
    int nl[4] = {0, 0, 0, 0};
    int ls[4][4] = {};
    int row[4][16] = {};
    int nlmax = 0;                                            // most degrees in one tile
    constexpr LmPack() {
        int fill[4] = {0, 0, 0, 0};
        for (int l = LMAX; l >= 0; --l) {
            int t = 0;
            while (fill[t] + 2 * l + 1 > 16) ++t;
            if (t + 1 > ntile) ntile = t + 1;
            ls[t][nl[t]++] = l;
            fill[t] += 2 * l + 1;
            if (nl[t] > nlmax) nlmax = nl[t];
        }
        for (int t = 0; t < 4; ++t) {
            int i = 0;
            for (int k = nl[t] - 1; k >= 0; --k)
                for (int m = -ls[t][k]; m <= ls[t][k]; ++m) row[t][i++] = ls[t][k] * ls[t][k] + ls[t][k] + m;
            for (; i < 16; ++i) row[t][i] = -1;
        }
    }
};

// sc T_{8q} .. sc T_{8q+7} of t (q = 0..3): the start pair (T_{8q}, T_{8q+1}) from the
// product rule T_{m+n} = 2 T_m T_n - T_{|m-n|}, then the three-term recurrence (fma)
__device__ __forceinline__ void cheb_rows8(float t, int q, float sc, float (&T)[8]) {
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
    // the recurrence is linear: scaling the start pair scales every row
    T[0] = a * sc;
    T[1] = b * sc;
#pragma unroll
    for (int k = 2; k < 8; ++k) T[k] = fmaf(2.0f * t, T[k - 1], -T[k - 2]);
}

// fp16 hi + lo of v (v - hi rounded again) as two h8 operands, built from packed conversions
// (two values per instruction, written straight into the operand dwords)
__device__ __forceinline__ void split8(const float (&v)[8], h8_t& hi, h8_t& lo) {
    union {
        h2_t p[4];
        h8_t v;
    } H, Lo;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const f2_t x = {v[2 * j], v[2 * j + 1]};
        const h2_t h = __builtin_convertvector(x, h2_t);
        H.p[j] = h;
        Lo.p[j] = __builtin_convertvector(x - __builtin_convertvector(h, f2_t), h2_t);
    }
    hi = H.v;
    lo = Lo.v;
}

template <int LMAX, int SB>
__global__ __launch_bounds__(256) void cheb_coeff_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                         const int32_t* __restrict__ atom_sp,
                                                         const double* __restrict__ shifts,
                                                         const double* __restrict__ centers,
                                                         const int32_t* __restrict__ self_sp,
                                                         const int64_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ nbr, int create_path, int n_centers,
                                                         int per_xcd, double inv_na, float tsc0, float tsc1,
                                                         float* __restrict__ out) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    constexpr LmPack<LMAX> kPack{};
    constexpr int LT = kPack.ntile;
    constexpr int NG = 2 * SB;                                // (species, interval) groups
    constexpr int TB = kCT * kCE, YB = (L2 + 1) * kCE;       // one operand buffer (halves)
    constexpr int GCAP = 4 * kCE + 8 * NG;                    // sorted positions, groups 8-aligned
    __shared__ __align__(16) _Float16 Th[2 * TB];             // two buffers: tile j and j + 1
    __shared__ __align__(16) _Float16 Tl[2 * TB];
    __shared__ __align__(16) _Float16 Yh[2 * YB];             // row L2 of a buffer: zeros
    __shared__ __align__(16) _Float16 Yl[2 * YB];
    __shared__ __align__(16) float4 geo[GCAP];                // (t, unit vector) by sorted position
    __shared__ int wcnt[4][NG];                               // edges per wavefront and group
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, cr = lane & 15;
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront
    // consecutive workgroups go to different XCDs: give each XCD a contiguous range of
    // centers, whose neighbourhoods overlap in its L2
    const int c = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (c >= n_centers) return;
    const int S = b.n_species, n = b.n_max;
    if (tid < 2 * kCE) {
        const int bb = tid >> 6;
        Yh[bb * YB + L2 * kCE + lane] = (_Float16)0.0f;
        Yl[bb * YB + L2 * kCE + lane] = (_Float16)0.0f;
    }
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int self_s = (create_path && self_sp) ? self_sp[c] : -1;
    const int cslot = self_sp[c] >= 0 ? self_sp[c] : S;
    const float rs = (float)b.cheb_split;
    // wavefront q owns the columns of lm tile q of LmPack (waves q >= LT only build
    // operands)
    const bool mine = q < LT;
    const int lt = q < LT ? q : 0;
    const int lm_col = kPack.row[lt][cr];                     // this lane's column (-1: none)
    const int yrow = lm_col >= 0 ? lm_col : L2;
    // this lane's column offset in rows r = 0..7 (mod 8) of an operand block (the XOR swizzle)
    int sw[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) sw[r] = ((((lane >> 3) ^ r) & 7) << 3) + (lane & 7);

    f4_t acc[NG][2];
#pragma unroll
    for (int gi = 0; gi < NG; ++gi)
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) acc[gi][kt] = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
    int ctot[NG];                                             // the center's edges per group
#pragma unroll
    for (int gi = 0; gi < NG; ++gi) ctot[gi] = 0;

    for (int64_t r0 = e0; r0 < e1; r0 += 4 * kCE) {
        // ---- geometry of a round of up to 256 edges (wavefront q, edges 64 q + lane),
        // sorted by (species, interval) over the whole round
        int grp = -1, rank = 0;
        float4 gv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        {
            const int64_t e = r0 + q * kCE + lane;
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
                    gv = make_float4(outer ? fmaf(r - rs, tsc1, -1.0f) : fmaf(r, tsc0, -1.0f), rx * ir, ry * ir,
                                     rz * ir);
                }
            }
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                const uint64_t m = __ballot(grp == gi);
                if (grp == gi)
                    rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (lane == 0) wcnt[q][gi] = __popcll(m);
            }
        }
        __syncthreads();
        // group gi holds positions [gs, gs + gn), its start aligned to 8, so that every lane's
        // 8 positions of an operand block belong to one group
        int gs[NG], gn[NG], nv = 0;
        {
            int before = 0;
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                const int c0 = wcnt[0][gi], c1 = wcnt[1][gi], c2 = wcnt[2][gi], c3 = wcnt[3][gi];
                gs[gi] = nv;
                gn[gi] = (c0 + c1) + (c2 + c3);
                ctot[gi] += gn[gi];
                if (grp == gi) before = nv + (q > 0 ? c0 : 0) + (q > 1 ? c1 : 0) + (q > 2 ? c2 : 0);
                nv += (gn[gi] + 7) & ~7;
            }
            if (grp >= 0) geo[before + rank] = gv;
        }
        __syncthreads();
        const int ntile = (nv + kCE - 1) / kCE;

        // ---- operands of tile j into buffer j & 1, lane = position 64 j + lane: wavefront q
        // writes the Chebyshev rows 8q .. 8q+7 and the harmonics of orders m = q (mod 4)
        auto build = [&](int j) {
            _Float16* th = Th + (j & 1) * TB;
            _Float16* tl = Tl + (j & 1) * TB;
            _Float16* yh = Yh + (j & 1) * YB;
            _Float16* yl = Yl + (j & 1) * YB;
            const int pp = kCE * j + lane;
            bool ok = false;
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) ok = ok || (pp >= gs[gi] && pp < gs[gi] + gn[gi]);
            const float4 u = geo[ok ? pp : 0];
            const float sc = ok ? kOpScale : 0.0f;
            float T[8];
            cheb_rows8(ok ? u.x : 0.0f, q, sc, T);
#pragma unroll
            for (int k = 0; k < 8; k += 2) {
                const int r0 = 8 * q + k;                     // rows r0, r0 + 1 (r0 = k mod 8)
                put_hilo2(th, tl, r0 * kCE + sw[k], (r0 + 1) * kCE + sw[k + 1], T[k], T[k + 1]);
            }
            harmonics_stripe_pairs<LMAX>(
                ok ? u.y : 0.0f, ok ? u.z : 0.0f, ok ? u.w : 0.0f, q, sc,
                [&](int l, float v) {
                    const int row = l * l + l;
                    put_hilo(yh, yl, row * kCE + sw[row & 7], v);
                },
                [&](int l, int m, float vp, float vm) {
                    const int rp = l * l + l + m, rm = l * l + l - m;
                    put_hilo2(yh, yl, rp * kCE + sw[rp & 7], rm * kCE + sw[rm & 7], vp, vm);
                });
        };
        // ---- products of tile j: wavefront q, columns lm = 16 q + lane % 16; load() issues the
        // LDS reads of one 32-position block, mma() its products
        struct Opnd {
            uint4 ah[2], al[2];
            h8_t bh, bl;
        };
        auto load = [&](int j, int blk, Opnd& o) {
            const _Float16* th = Th + (j & 1) * TB;
            const _Float16* tl = Tl + (j & 1) * TB;
            const _Float16* yh = Yh + (j & 1) * YB;
            const _Float16* yl = Yl + (j & 1) * YB;
            const int ch = 4 * blk + g;
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) {
                const int row = 16 * kt + cr;
                const int off = row * kCE + (((ch ^ row) & 7) << 3);
                o.ah[kt] = *(const uint4*)(th + off);
                o.al[kt] = *(const uint4*)(tl + off);
            }
            const int off = yrow * kCE + (((ch ^ yrow) & 7) << 3);
            o.bh = *(const h8_t*)(yh + off);
            o.bl = *(const h8_t*)(yl + off);
        };
        auto mma = [&](int j, int blk, const Opnd& o) {
            const int b0 = kCE * j + 32 * blk, cs = b0 + 8 * g;   // this lane's 8 positions: [cs, cs + 8)
            const h8_t x0 = __builtin_bit_cast(h8_t, o.ah[0]), x1 = __builtin_bit_cast(h8_t, o.ah[1]);
            const h8_t y0 = __builtin_bit_cast(h8_t, o.al[0]), y1 = __builtin_bit_cast(h8_t, o.al[1]);
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                const int lo = gs[gi], hi = gs[gi] + ((gn[gi] + 7) & ~7);
                if (gn[gi] == 0 || hi <= b0 || lo >= b0 + 32) continue;
                h8_t bh = o.bh, bl = o.bl;
                if (!(lo <= b0 && hi >= b0 + 32)) {
                    const bool in = cs >= lo && cs < hi;
                    const h8_t z = h8_t{};
                    bh = in ? bh : z;
                    bl = in ? bl : z;
                }
                acc[gi][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(x0, bh, acc[gi][0], 0, 0, 0);
                acc[gi][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(x1, bh, acc[gi][1], 0, 0, 0);
                acc[gi][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(x0, bl, acc[gi][0], 0, 0, 0);
                acc[gi][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(x1, bl, acc[gi][1], 0, 0, 0);
                acc[gi][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(y0, bh, acc[gi][0], 0, 0, 0);
                acc[gi][1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(y1, bh, acc[gi][1], 0, 0, 0);
            }
        };
        build(0);
        __syncthreads();
        for (int j = 0; j < ntile; ++j) {
            // the first block's reads of tile j are in flight while this wavefront writes the
            // operands of tile j + 1 (other buffer); its products follow
            const bool two = nv - kCE * j > 32;
            Opnd o;
            if (mine) load(j, 0, o);
            if (j + 1 < ntile) build(j + 1);
            if (mine) {
                mma(j, 0, o);
                if (two) {
                    load(j, 1, o);
                    mma(j, 1, o);
                }
            }
            __syncthreads();
        }
    }
    if (!mine) return;

    // ---- c[s][n][lm] = sum_{p,k} a_lnpk Q_sp[k][lm] on the matrix cores: A = the fp16 hi/lo
    // coefficient rows (n) of each degree of this wavefront's tile, B = Q from the
    // accumulators (this lane: column lm, k = 4 g + (j % 4) + 16 (j / 4)), columns of other
    // degrees zeroed; B is scaled by 2^e_q (|Q| <= 1.03 x the group's edge count) before the
    // hi/lo split.  All fragments of a species are loaded before its products.
    const int dl = lm_col >= 0 ? degree_of(lm_col, LMAX) : -1;
    constexpr int NLW = kPack.nlmax;                          // degrees per tile, at most
    const int nlw = kPack.nl[lt];
    const h8_t* frag = (const h8_t*)b.cheb_frag;
#pragma unroll
    for (int s = 0; s < SB; ++s) {
        if (s >= S) break;
        const int slot = cslot * S + s;
        h8_t fa[NLW][2][2];
#pragma unroll
        for (int li = 0; li < NLW; ++li) {
            if (li >= nlw) break;
            const int l = kPack.ls[lt][li];
#pragma unroll
            for (int pp = 0; pp < 2; ++pp) {
                const h8_t* fr = frag + (((size_t)slot * L1 + l) * 2 + pp) * 64 * 2 + 2 * lane;
                fa[li][pp][0] = fr[0];
                fa[li][pp][1] = fr[1];
            }
        }
        const int nmax_g = max(ctot[2 * s], ctot[2 * s + 1]);
        // 2^eq (1.03 nmax_g) <= 2^14
        const int eq = 14 - (32 - __builtin_clz((uint32_t)(nmax_g + (nmax_g >> 5) + 1)));
        const float qs = ldexpf(kOpUnscale, eq);
        h8_t bh[2], bl[2];
#pragma unroll
        for (int pp = 0; pp < 2; ++pp) {
            float v[8];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) v[jj] = acc[2 * s + pp][jj >> 2][jj & 3] * qs;
            split8(v, bh[pp], bl[pp]);
        }
        // one accumulator per degree of the tile (every column, rows n of that degree's
        // coefficients); each lane keeps the one of its column's degree
        f4_t cacc = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int li = 0; li < NLW; ++li) {
            if (li >= nlw) break;
            f4_t dacc = f4_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int pp = 0; pp < 2; ++pp) {
                dacc = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[li][pp][0], bh[pp], dacc, 0, 0, 0);
                dacc = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[li][pp][0], bl[pp], dacc, 0, 0, 0);
                dacc = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[li][pp][1], bh[pp], dacc, 0, 0, 0);
            }
            if (dl == kPack.ls[lt][li]) cacc = dacc;
        }
        // this lane: rows n = 4 g + i of column lm
        if (g < 2 && lm_col >= 0) {
            const float* rinv = b.cheb_rowinv + ((size_t)slot * L1 + dl) * 8;
            const float us = ldexpf(1.0f, -eq);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int nn = 4 * g + i;
                if (nn >= n) continue;
                float v = cacc[i] * rinv[nn] * us;
                if (s == self_s && lm_col == 0) v += (float)b.self_l0[nn];
                out[(((size_t)c * S + s) * n + nn) * L2 + lm_col] = v;
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
        // 1 / n_atoms (the extended-index split) and the interval scales 2 / width, once
        const double inv_na = 1.0 / (double)n_atoms;
        const float tsc0 = (float)(2.0 / b.cheb_split), tsc1 = (float)(2.0 / (b.cutoff - b.cheb_split));
        cheb_coeff_kernel<LMAX, SB><<<8 * per, 256, 0, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, self_sp,
                                                             offsets, nbr, create_path, n_centers, per, inv_na, tsc0,
                                                             tsc1, out);
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
    if (!b->cheb_frag || !b->cheb_rowinv)
        return fail(CUSOAP_ERR_ARG, "coefficients: the Chebyshev form needs cheb_frag and cheb_rowinv");
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
