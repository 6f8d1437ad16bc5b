// Fused density-coefficient kernel c[center][s][n][lm].
//
// Restates _coefficients_gto / _coefficients_polynomial / _add_self_terms
// (soap_generator.py:1294-1435, 821-850) and the Cacc accumulation of the
// closed-form derivative path (soap_generator.py:2414-2423) as one pass per
// center: one 256-thread workgroup per center walks the center's CSR edge
// segment in tiles of 64 edges.
//
// Phase A (one pass, no barrier inside): in every wavefront lane e holds edge e
// of the tile and recomputes its geometry (the repeated gathers hit the cache).
// Wavefront q evaluates the harmonics of orders m = q (mod 4) and the radial
// functions of degrees l = q (mod 4) (polynomial basis: radial index k = q
// mod 4), so the four wavefronts split each edge's work without divergence.
// Results go to LDS rows per edge: harmonics in groups of 4 aligned to 16 bytes
// (one group = 4 consecutive m of one l), radial values [l][n] with n padded to
// even.  Every wavefront also holds, from ballots, a 64-bit mask per species
// of the tile's edges that count.
// The tile's counted edges are compacted by species (ballot + mbcnt ranks), and
// the LDS rows are indexed by that position, so every loop below runs over a
// contiguous, wavefront-uniform range.
// Phase B: lane (n pair, lm group) of wavefront w accumulates (2 n) x (4 lm) x
// (species) for the positions p = w (mod 4) of each species range (no per-edge
// branch, no wasted products), one 8-byte radial and one 16-byte harmonic LDS
// read per edge.
// The four wavefronts' partial sums are combined in LDS in a fixed order, so
// results are bitwise reproducible.
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
// float accumulation except for F64.  The table is read with the lanes of a
// wavefront on the (l, n) channels of one edge (coalesced node rows), not one
// edge per lane, which would scatter every lane over its own rows.
// GTO constants (gamma, pref, beta) are staged in LDS once per workgroup.
//
// create_path: edges with |r| <= 1e-12 are skipped and the l=0 self term is
// added for index centers, as the reference does; otherwise every edge counts.
#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kRbfTable = 2;     // kernel variant: tabulated radial model
constexpr int kTE = 64;          // edges per tile (one per lane)
constexpr int kThreads = 256;

__host__ __device__ constexpr int n_groups(int lmax) {
    int g = 0;
    for (int l = 0; l <= lmax; ++l) g += (2 * l + 1 + 3) / 4;
    return g;
}

// first lm group of degree l
__host__ __device__ constexpr int group_start(int l) {
    int g = 0;
    for (int k = 0; k < l; ++k) g += (2 * k + 1 + 3) / 4;
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

// LDS constants of the GTO basis, rounded to 16 bytes
__host__ __device__ inline int gcon_doubles(bool gto, int L1, int n) { return gto ? ((L1 * n * (2 + n) + 1) & ~1) : 0; }

// LDS row strides (elements): harmonics 4 NG + 4, radial (l_max+1) n_even + 2
__host__ __device__ inline int y_stride(int lmax) { return 4 * n_groups(lmax) + 4; }
__host__ __device__ inline int r_stride(int lmax, int n) { return (lmax + 1) * ((n + 1) & ~1) + 2; }

template <typename T>
struct V2;
template <>
struct V2<float> {
    using type = float2;
};
template <>
struct V2<double> {
    using type = double2;
};

// 4 wavefronts per SIMD for the float accumulators (a few spilled registers cost
// less than the lost occupancy); 3 for float64, which would spill more
template <int PREC>
constexpr int coeff_min_waves() { return PREC == CUSOAP_PREC_F64 ? 3 : 4; }

// SB: species accumulated per pass over the edges (2 when the system has at most 2,
// else 4): the accumulators of unused species would otherwise hold registers and
// push the kernel into spills at 4 wavefronts per SIMD
template <int LMAX, int PREC, int RBF, int SB>
__global__ __launch_bounds__(kThreads, coeff_min_waves<PREC>()) void coeff_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
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
    constexpr bool kGtoConst = RBF == CUSOAP_RBF_GTO;
    constexpr int L1 = LMAX + 1, L2 = L1 * L1, NG = n_groups(LMAX);
    const int n = b.n_max, S = b.n_species;
    const int nP = (n + 1) & ~1;
    const int YS = y_stride(LMAX), RS = r_stride(LMAX, n);
    extern __shared__ __align__(16) unsigned char smem[];
    double* gcon = (double*)smem;                     // gamma[L1 n], pref[L1 n], beta[L1 n n] (GTO)
    TA* tab_w = (TA*)(gcon + gcon_doubles(kGtoConst, L1, n));               // [4 waves][16][4] weights (table)
    int32_t* tab_row = (int32_t*)(tab_w + (RBF == kRbfTable ? 4 * kTE : 0)); // [4 waves][16] row offsets (table)
    TY* Ys = (TY*)(tab_row + (RBF == kRbfTable ? kTE : 0));                 // [kTE][YS]
    TA* Rs = (TA*)(Ys + kTE * YS);                    // [kTE][RS]; later the combine / beta stage

    const int c = blockIdx.x, tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const int q = __builtin_amdgcn_readfirstlane(wv);
    if (kGtoConst) {
        const int nc = L1 * n;
        for (int i = tid; i < nc; i += kThreads) {
            gcon[i] = b.gto_gamma[i];
            gcon[nc + i] = b.gto_pref[i];
        }
        if (PREC == CUSOAP_PREC_MIXED)
            for (int i = tid; i < nc * n; i += kThreads) gcon[2 * nc + i] = b.gto_beta[i];
    }
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int npair = (n + 1) / 2;
    const int nslot_all = npair * NG;                 // (n pair, lm group) slots per species
    // few slots (small n, l): the wavefront splits into `split` lane groups of `width`
    // lanes, group h summing every split-th of the wavefront's positions (offset h);
    // the groups' sums are added with cross-lane shuffles before the combine
    // (l_max <= 4 only, as for the channel pairs)
    const int split = LMAX > 4 ? 1 : (nslot_all <= 16 ? 4 : (nslot_all <= 32 ? 2 : 1));
    const int width = 64 / split, lgrp = lane / width;
    const int self_s = (create_path && self_sp) ? self_sp[c] : -1;
    const int cslot = (RBF == kRbfTable && self_sp[c] >= 0) ? self_sp[c] : S;   // table slot of the center
    const double* g_gamma = gcon;
    const double* g_pref = gcon + L1 * n;
    const double* g_beta = gcon + 2 * L1 * n;
    // table phase: this lane's (l, n) channel of the first 64; grid constants
    const int L1n = L1 * n, ch_l = lane / n, ch_n = lane - (lane / n) * n;
    // at most 32 channels (small n, l): the two half-wavefronts take two edge positions
    // at a time, lane & 31 being the channel
    // (compiled in for l_max <= 4 only: larger l_max keeps the plain code)
    const bool ch_pair = LMAX <= 4 && L1n <= 32;
    const int pc = lane & 31, pc_l = pc / n, pc_n = pc - (pc / n) * n, phalf = lane >> 5;
    const double tab_inv_h = RBF == kRbfTable ? b.rad_intervals / b.rad_rmax : 0.0;
    const TA tab_h = RBF == kRbfTable ? (TA)(b.rad_rmax / b.rad_intervals) : TA(0);
    // every row slot holds a valid table offset from the start, so the loads of slots
    // past a wavefront's positions (whose results phase B never reads) stay in bounds
    if (RBF == kRbfTable)
        for (int i = tid; i < kTE; i += kThreads) tab_row[i] = 0;
    const double inv_na = 1.0 / (double)n_atoms;
    __syncthreads();

    for (int s0 = 0; s0 < S; s0 += SB) {
        const int s1 = (s0 + SB < S) ? s0 + SB : S;
        for (int q0 = 0; q0 < nslot_all; q0 += width) {
            // ---- this lane's slot: n pair p (n0 = 2p, n1 = 2p+1) x lm group g of degree l
            const int qs = q0 + (lane - lgrp * width);
            const bool valid = qs < nslot_all;
            const bool owner = valid && lgrp == 0;          // writes the slot's combined sum
            int n0 = 0, l = 0, lm0 = 0, cnt = 0, grp = 0;
            bool has1 = false;
            if (valid) {
                const int pp = qs / NG;
                grp = qs - pp * NG;
                int g = grp;
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
            const int roff = l * nP + n0, yoff = 4 * grp;
            TA acc[SB][2][4];
#pragma unroll
            for (int a4 = 0; a4 < SB; ++a4)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[a4][h][j] = TA(0);

            // neighbour indices are loaded one tile ahead, so the first load of the
            // nbr -> position chain is off each tile's critical path
            int ext_next = (e0 + lane < e1) ? nbr[e0 + lane] : 0;
            for (int64_t t0 = e0; t0 < e1; t0 += kTE) {
                const int nt = (int)((e1 - t0) < kTE ? (e1 - t0) : kTE);
                const int ext_cur = ext_next;
                if (t0 + kTE + lane < e1) ext_next = nbr[t0 + kTE + lane];
                // ---- A: lane = edge; wavefront q takes m = q, l = q (mod 4)
                int sp = -1;
                double rx = 0.0, ry = 0.0, rz = 0.0, r2 = 0.0, r = 0.0, w = 1.0;
                if (lane < nt) {
                    int img, a;
                    split_ext(ext_cur, n_atoms, inv_na, img, a);
                    rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
                    ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
                    rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
                    r2 = (rx * rx + ry * ry) + rz * rz;
                    r = sqrt(r2);
                    if (b.weighting != CUSOAP_W_NONE) {
                        double dw = 0.0;
                        weighting(b, r, w, dw);
                    }
                    const int spa = atom_sp[a];
                    sp = ((create_path && !(r > 1e-12)) || spa < s0 || spa >= s1) ? -1 : spa;
                }
                // compact the tile's counted edges by species: position p = start of the
                // species + rank of the edge among that species' edges (ballot + mbcnt), the
                // same in every wavefront; LDS rows are indexed by p, so the loops below run
                // over contiguous uniform ranges instead of walking lane masks
                int beg[SB], scnt[SB], pidx = -1, nv = 0;
#pragma unroll
                for (int a4 = 0; a4 < SB; ++a4) {
                    const uint64_t m = __ballot(sp == s0 + a4);
                    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    beg[a4] = nv;
                    scnt[a4] = __popcll(m);
                    if (sp == s0 + a4) pidx = nv + rank;
                    nv += scnt[a4];
                }
                if (RBF == kRbfTable) {
                    // radial table: lanes = (l, n) channels of one edge, so each edge reads its
                    // two node rows with coalesced 8-byte (value, slope) loads; wavefront q takes
                    // positions p = q + 4k (k < 16), four per round so that their loads are in
                    // flight together.  The edge's own lane first stores its weighted Hermite
                    // factors and row offset in the wavefront's slots k.
                    const int G = b.rad_intervals, NS = L1 * 2 * n;
                    TA* wq = tab_w + q * 4 * 16;
                    int32_t* rq = tab_row + q * 16;
                    if (pidx >= 0 && (pidx & 3) == q) {
                        const double xr = r * tab_inv_h;
                        int i = (int)xr;
                        i = i < G - 1 ? i : G - 1;
                        const TA t = (TA)(xr - i), tw = (TA)w;
                        const TA t2 = t * t, t3 = t2 * t;
                        TA* hw = wq + (pidx >> 2) * 4;
                        hw[0] = tw * (TA(2) * t3 - TA(3) * t2 + TA(1));
                        hw[1] = tw * tab_h * (t3 - TA(2) * t2 + t);
                        hw[2] = tw * (TA(3) * t2 - TA(2) * t3);
                        hw[3] = tw * tab_h * (t3 - t2);
                        rq[pidx >> 2] = (int32_t)(4u * (uint32_t)(((cslot * S + sp) * (G + 1) + i) * NS));   // bytes
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const int nk = (nv - q + 3) >> 2;       // this wavefront's positions
                    // loads address the table as its uniform base plus a 32-bit byte offset
                    const char* tb = (const char*)b.rad_table;
                    const uint32_t ns_b = 4u * (uint32_t)NS;
                    if (ch_pair) {
                        const bool on = pc < L1n;
                        const uint32_t coff = on ? 8u * (uint32_t)(pc_l * n + pc_n) : 0u;
                        TA* rb = Rs + q * RS + pc_l * nP + pc_n;
                        for (int k0 = 0; k0 < nk; k0 += 4) {
                            // half h: slots j = h and j = h + 2
                            const int4 rows = *(const int4*)(rq + k0);
                            const uint32_t oa = (uint32_t)(phalf ? rows.y : rows.x) + coff;
                            const uint32_t ob = (uint32_t)(phalf ? rows.w : rows.z) + coff;
                            const float2 va0 = *(const float2*)(tb + oa), va1 = *(const float2*)(tb + (oa + ns_b));
                            const float2 vb0 = *(const float2*)(tb + ob), vb1 = *(const float2*)(tb + (ob + ns_b));
                            if (on) {
                                const int ja = k0 + phalf, jb = k0 + phalf + 2;
                                const typename V2<TA>::type wa01 = *(const typename V2<TA>::type*)(wq + ja * 4);
                                const typename V2<TA>::type wa23 = *(const typename V2<TA>::type*)(wq + ja * 4 + 2);
                                const typename V2<TA>::type wb01 = *(const typename V2<TA>::type*)(wq + jb * 4);
                                const typename V2<TA>::type wb23 = *(const typename V2<TA>::type*)(wq + jb * 4 + 2);
                                rb[ja * 4 * RS] = wa01.x * (TA)va0.x + wa01.y * (TA)va0.y +
                                                          wa23.x * (TA)va1.x + wa23.y * (TA)va1.y;
                                rb[jb * 4 * RS] = wb01.x * (TA)vb0.x + wb01.y * (TA)vb0.y +
                                                          wb23.x * (TA)vb1.x + wb23.y * (TA)vb1.y;
                            }
                        }
                    } else {
                        for (int ch0 = 0; ch0 < L1n; ch0 += 64) {
                            const bool on = ch0 + lane < L1n;
                            int ll = ch_l, nn = ch_n;
                            if (ch0 > 0) {
                                ll = (ch0 + lane) / n;
                                nn = ch0 + lane - ll * n;
                            }
                            const uint32_t coff = on ? 8u * (uint32_t)(ll * n + nn) : 0u;
                            TA* rb = Rs + q * RS + ll * nP + nn;
                            for (int k0 = 0; k0 < nk; k0 += 4) {
                                const int4 rows = *(const int4*)(rq + k0);
                                const uint32_t o[4] = {(uint32_t)rows.x + coff, (uint32_t)rows.y + coff,
                                                       (uint32_t)rows.z + coff, (uint32_t)rows.w + coff};
                                float2 v0[4], v1[4];
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    v0[j] = *(const float2*)(tb + o[j]);
                                    v1[j] = *(const float2*)(tb + (o[j] + ns_b));
                                }
                                // slots j >= ne (k0 + j < 16, so position < 64) fill rows that phase B
                                // never reads: no per-slot branch
                                if (on) {
                                    TA* rr = rb + k0 * 4 * RS;
#pragma unroll
                                    for (int j = 0; j < 4; ++j) {
                                        const typename V2<TA>::type hw01 = *(const typename V2<TA>::type*)(wq + (k0 + j) * 4);
                                        const typename V2<TA>::type hw23 = *(const typename V2<TA>::type*)(wq + (k0 + j) * 4 + 2);
                                        rr[j * 4 * RS] = hw01.x * (TA)v0[j].x + hw01.y * (TA)v0[j].y +
                                                                 hw23.x * (TA)v1[j].x + hw23.y * (TA)v1[j].y;
                                    }
                                }
                            }
                        }
                    }
                }
                if (pidx >= 0) {
                    TY* yrow = Ys + pidx * YS;
                    solid_harmonics_stripe<LMAX, TY>((TY)rx, (TY)ry, (TY)rz, q, [&](int ll, int m, TY v) {
                        yrow[4 * group_start(ll) + ll + m] = v;
                    });
                    TA* rrow = Rs + pidx * RS;
                    if (RBF == kRbfTable) {
                        // radial values written above
                    } else if (RBF == CUSOAP_RBF_POLY) {
                        for (int k = q; k < n; k += 4) {
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
                            for (int ll = 0; ll < L1; ++ll) rrow[ll * nP + k] = (TA)(w * I[ll]);
                        }
                    } else {
#pragma unroll
                        for (int ll = 0; ll < L1; ++ll) {
                            if ((ll & 3) != q) continue;
                            if (kBetaPerCenter) {
                                for (int k = 0; k < n; ++k) {
                                    const TQ pref = (TQ)g_pref[ll * n + k], gam = (TQ)g_gamma[ll * n + k];
                                    rrow[ll * nP + k] = (TA)((TQ)w * (pref * t_exp<TQ>(-gam * (TQ)r2)));
                                }
                            } else {
                                double qv[kMaxN];
#pragma unroll
                                for (int k = 0; k < kMaxN; ++k)
                                    qv[k] = k < n ? g_pref[ll * n + k] * exp(-g_gamma[ll * n + k] * r2) : 0.0;
                                for (int nn = 0; nn < n; ++nn) {
                                    const double* beta = g_beta + ((size_t)ll * n + nn) * n;
                                    double B = 0.0;
#pragma unroll
                                    for (int k = 0; k < kMaxN; ++k)
                                        if (k < n) B += beta[k] * qv[k];
                                    rrow[ll * nP + nn] = (TA)(w * B);
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                // ---- B: wavefront q sums the positions p = q (mod 4) of each species range
                if (valid) {
#pragma unroll
                    for (int a4 = 0; a4 < SB; ++a4) {
                        if (s0 + a4 >= s1) continue;
                        auto fma_edge = [&](int e) {
                            const typename V2<TA>::type rv = *(const typename V2<TA>::type*)(Rs + e * RS + roff);
                            const typename V2<TY>::type y01 = *(const typename V2<TY>::type*)(Ys + e * YS + yoff);
                            const typename V2<TY>::type y23 = *(const typename V2<TY>::type*)(Ys + e * YS + yoff + 2);
                            const TA yv[4] = {(TA)y01.x, (TA)y01.y, (TA)y23.x, (TA)y23.y};
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                acc[a4][0][j] += rv.x * yv[j];
                                acc[a4][1][j] += rv.y * yv[j];
                            }
                        };
                        // two positions per step so that their LDS reads are in flight together
                        const int pe = beg[a4] + scnt[a4], ps = 4 * split;
                        int pp = beg[a4] + ((q - beg[a4]) & 3) + 4 * lgrp;
                        for (; pp + ps < pe; pp += 2 * ps) {
                            fma_edge(pp);
                            fma_edge(pp + ps);
                        }
                        if (pp < pe) fma_edge(pp);
                    }
                }
                __syncthreads();
            }
            // ---- lane groups of one slot: fixed-order butterfly (every lane of the wavefront
            // takes part; lanes of invalid slots carry zeros)
            for (int o = 32; o >= width; o >>= 1) {
#pragma unroll
                for (int a4 = 0; a4 < SB; ++a4)
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[a4][h][j] += __shfl_xor(acc[a4][h][j], o, 64);
            }
            // ---- combine the four wavefronts' partial sums in a fixed order (LDS, aliases the tile).
            // When four [SB][n][L2] blocks fit in the tile region, each wavefront stores its own
            // block and one barrier precedes the sum ((w0 + w1) + w2) + w3; otherwise the
            // wavefronts add into one block in turn (the same order, four barriers).
            TA* part = Rs;                                  // [SB species][n][L2]
            const int psz = SB * n * L2;
            const int region = (kTE * RS > (S > 4 ? S : 4) * n * L2) ? kTE * RS : (S > 4 ? S : 4) * n * L2;
            const bool blocks4 = 4 * psz <= region;
            if (blocks4) {
                if (owner) {
                    TA* mine = part + wv * psz;
#pragma unroll
                    for (int a4 = 0; a4 < SB; ++a4) {
                        if (s0 + a4 >= s1) continue;
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            if (h == 1 && !has1) continue;
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                if (j >= cnt) continue;
                                mine[(a4 * n + n0 + h) * L2 + lm0 + j] = acc[a4][h][j];
                            }
                        }
                    }
                }
                __syncthreads();
                if (wv == 0 && owner) {
                    for (int a4 = 0; a4 < s1 - s0; ++a4)
                        for (int h = 0; h < 2; ++h) {
                            if (h == 1 && !has1) continue;
                            const int nn = n0 + h;
                            TA* dst = out + (((size_t)c * S + s0 + a4) * n + nn) * L2 + lm0;
                            for (int j = 0; j < cnt; ++j) {
                                const int o = (a4 * n + nn) * L2 + lm0 + j;
                                TA v = ((part[o] + part[psz + o]) + part[2 * psz + o]) + part[3 * psz + o];
                                if (!kBetaPerCenter && s0 + a4 == self_s && lm0 + j == 0) v += (TA)b.self_l0[nn];
                                dst[j] = v;
                            }
                        }
                }
                __syncthreads();
                continue;
            }
            for (int w2 = 0; w2 < 4; ++w2) {
                if (wv == w2 && owner) {
#pragma unroll
                    for (int a4 = 0; a4 < SB; ++a4) {
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
            if (wv == 0 && owner) {
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
    const int n = b.n_max;
    const size_t ys_bytes = sizeof(TY) * (size_t)kTE * y_stride(LMAX);
    const size_t rs_bytes = sizeof(TA) * (size_t)kTE * r_stride(LMAX, n);
    const size_t part = sizeof(TA) * (size_t)std::max(b.n_species, 4) * n * L2;   // combine / beta stage (aliases Rs)
    const size_t gcon = sizeof(double) * (size_t)gcon_doubles(RBF == CUSOAP_RBF_GTO, L1, n);
    const size_t tab = RBF == kRbfTable ? (sizeof(int32_t) + 4 * sizeof(TA)) * (size_t)kTE : 0;
    const size_t smem = gcon + tab + ys_bytes + (part > rs_bytes ? part : rs_bytes);
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: LDS tile too large");
    auto kern = b.n_species <= 2 ? coeff_kernel<LMAX, PREC, RBF, 2> : coeff_kernel<LMAX, PREC, RBF, 4>;
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

// build.py compiles this file once per l_max value with -DCUSOAP_ONLY_L=<l> (that object
// instantiates the kernels of one degree, so the degrees compile in parallel) and once
// without it, for the entry point below, which calls the per-degree functions.
#define CUSOAP_COEFF_PARAMS const cusoap_basis_t& b, int prec, const double* pos, int n_atoms, \
    const int32_t* atom_sp, const double* shifts, const double* centers, int n_centers, const int32_t* self_sp, \
    const int64_t* offsets, const int32_t* nbr, int create_path, void* out, hipStream_t st
#define CUSOAP_COEFF_DECL(L) int coefficients_l##L(CUSOAP_COEFF_PARAMS);
CUSOAP_COEFF_DECL(0) CUSOAP_COEFF_DECL(1) CUSOAP_COEFF_DECL(2) CUSOAP_COEFF_DECL(3) CUSOAP_COEFF_DECL(4)
CUSOAP_COEFF_DECL(5) CUSOAP_COEFF_DECL(6) CUSOAP_COEFF_DECL(7) CUSOAP_COEFF_DECL(8) CUSOAP_COEFF_DECL(9)
#undef CUSOAP_COEFF_DECL

#ifdef CUSOAP_ONLY_L
#define CUSOAP_COEFF_DEF(L)                                                                                    \
    int coefficients_l##L(CUSOAP_COEFF_PARAMS) {                                                             \
        return dispatch_prec<L>(b, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, \
                                create_path, out, st);                                                       \
    }
#define CUSOAP_COEFF_DEF2(L) CUSOAP_COEFF_DEF(L)
CUSOAP_COEFF_DEF2(CUSOAP_ONLY_L)
#else
int cheb_coefficients_impl(const cusoap_basis_t* b, const double* pos, int n_atoms, const int32_t* atom_sp,
                           const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                           const int64_t* offsets, const int32_t* nbr, int create_path, float* out, hipStream_t st);

int coefficients_impl(const cusoap_basis_t* b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                      const double* shifts, int n_img, const double* centers, int n_centers,
                      const int32_t* self_sp, const int64_t* offsets, const int32_t* nbr, int create_path,
                      void* out, hipStream_t st) {
    (void)n_img;
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL || b->n_species < 1)
        return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: n_max must be 1..16 and l_max 0..9");
    if (prec < 0 || prec > 2) return fail(CUSOAP_ERR_ARG, "coefficients: unknown precision");
    if (create_path && !self_sp) return fail(CUSOAP_ERR_ARG, "coefficients: create path needs center_self_species");
    if (b->rad_cheb) {
        if (prec == CUSOAP_PREC_F64)
            return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: the Chebyshev radial form computes in float32");
        return cheb_coefficients_impl(b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr,
                                      create_path, (float*)out, st);
    }
    if (b->rad_table && (!self_sp || b->rad_intervals < 1 || !(b->rad_rmax > 0.0)))
        return fail(CUSOAP_ERR_ARG, "coefficients: radial table needs center_self_species and a valid grid");
    if (b->rad_table && (double)(b->n_species + 1) * b->n_species * (b->rad_intervals + 1.0) * (b->l_max + 1) * 2 * b->n_max >= 1073741824.0)
        return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: radial table larger than 2^30 values");
    if (n_centers == 0) return CUSOAP_OK;
#define CUSOAP_CASE(L) \
    case L: return coefficients_l##L(*b, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st);
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: break;
    }
#undef CUSOAP_CASE
    return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: l_max out of range");
}
#endif  // CUSOAP_ONLY_L

}  // namespace cusoap
