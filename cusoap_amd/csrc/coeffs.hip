// Fused density-coefficient kernel c[center][s][n][lm].
//
// Restates _coefficients_gto / _coefficients_polynomial / _add_self_terms
// (soap_generator.py:1294-1435, 821-850) and the Cacc accumulation of the
// closed-form derivative path (soap_generator.py:2414-2423) as one pass:
//
//   one wavefront (64 lanes) per center; for each species s in turn the
//   center's CSR edge segment is stream-compacted (ballot + popcount) into
//   LDS tiles of 64 edges of that species.  Per tile, lane e evaluates its
//   edge: real solid harmonics r^l Y_lm (angular type TY) and the radial
//   factors (type TR) -- GTO primitives w * pref_kl * exp(-gamma_kl r^2), or
//   the orthonormal polynomial-basis B_nl = w I_nl(r) / r^l -- into LDS.  The
//   lanes then own (n, lm) accumulator slots in registers and sweep the tile.
//   GTO accumulators hold primitive sums; the Loewdin beta_l contraction is
//   applied once per center (S n^2 L^2 work instead of per edge).
//
// The center's own atom (create path) is not an edge; its l=0 self term is
// added at the end, as the reference does.
#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kTile = 64;

template <int LMAX>
struct Dims {
    static constexpr int L1 = LMAX + 1;
    static constexpr int L2 = L1 * L1;
    static constexpr int NSLOT = (kMaxN * L2 + 63) / 64;
};

__device__ __forceinline__ int lm_to_l(int lm) {
    int l = (int)sqrtf((float)lm);
    while (l * l > lm) --l;
    while ((l + 1) * (l + 1) <= lm) ++l;
    return l;
}

template <typename T>
__device__ __forceinline__ T t_exp(T v);
template <>
__device__ __forceinline__ float t_exp<float>(float v) { return expf(v); }
template <>
__device__ __forceinline__ double t_exp<double>(double v) { return exp(v); }

template <int LMAX, typename TY, typename TR, int RBF>
__global__ __launch_bounds__(64) void coeff_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                   const int32_t* __restrict__ atom_sp,
                                                   const double* __restrict__ shifts,
                                                   const double* __restrict__ centers,
                                                   const int32_t* __restrict__ self_sp,
                                                   const int64_t* __restrict__ offsets,
                                                   const int32_t* __restrict__ nbr, int create_path,
                                                   TR* __restrict__ out) {
    using D = Dims<LMAX>;
    constexpr int L1 = D::L1, L2 = D::L2, NSLOT = D::NSLOT;
    const int n = b.n_max, S = b.n_species;
    const int NL = n * L1;
    extern __shared__ __align__(16) unsigned char smem[];
    TR* Rsh = (TR*)smem;                              // [kTile][NL]
    TR* accsh = Rsh + kTile * NL;                     // [n][L2]
    TY* Ysh = (TY*)(accsh + n * L2);                  // [kTile][L2]
    int32_t* esh = (int32_t*)(Ysh + kTile * L2);      // [2 * kTile] compacted ext indices

    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const double cx = centers[3 * (size_t)c], cy = centers[3 * (size_t)c + 1], cz = centers[3 * (size_t)c + 2];
    const int64_t e0 = offsets[c], e1 = offsets[c + 1];
    const int nslots = n * L2;

    int slot_r[NSLOT], slot_y[NSLOT];
#pragma unroll
    for (int i = 0; i < NSLOT; ++i) {
        const int t = lane + 64 * i;
        const int nn = t / L2, lm = t - nn * L2;
        slot_y[i] = lm;
        slot_r[i] = nn * L1 + lm_to_l(lm);
    }

    for (int s = 0; s < S; ++s) {
        TR acc[NSLOT];
#pragma unroll
        for (int i = 0; i < NSLOT; ++i) acc[i] = TR(0);
        int64_t cursor = e0;
        int cnt = 0;
        while (cursor < e1 || cnt > 0) {
            // ---- gather up to one tile of species-s edges
            while (cnt < kTile && cursor < e1) {
                const int64_t e = cursor + lane;
                int ext = 0;
                bool match = false;
                if (e < e1) {
                    ext = nbr[e];
                    match = atom_sp[ext % n_atoms] == s;
                }
                const uint64_t bal = __ballot(match);
                const int pre = __popcll(bal & ((1ull << lane) - 1ull));
                if (match) esh[cnt + pre] = ext;
                cnt += __popcll(bal);
                cursor += 64;
                __syncthreads();
            }
            const int ntile = cnt < kTile ? cnt : kTile;
            // ---- evaluate the tile's edges
            if (lane < ntile) {
                const int ext = esh[lane];
                const int img = ext / n_atoms, a = ext - img * n_atoms;
                const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
                const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
                const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
                const double r2 = (rx * rx + ry * ry) + rz * rz;
                const double r = sqrt(r2);
                double w = 1.0, dw = 0.0;
                if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
                TY* Yrow = Ysh + lane * L2;
                solid_harmonics<LMAX, false, TY>((TY)rx, (TY)ry, (TY)rz,
                                                 [&](int lm, TY v, TY, TY, TY) { Yrow[lm] = v; });
                TR* Rrow = Rsh + lane * NL;
                if (RBF == CUSOAP_RBF_GTO) {
                    const TR tr2 = (TR)r2, tw = (TR)w;
                    for (int k = 0; k < n; ++k) {
#pragma unroll
                        for (int l = 0; l < L1; ++l) {
                            const TR pref = (TR)b.gto_pref[l * n + k];
                            const TR gam = (TR)b.gto_gamma[l * n + k];
                            Rrow[k * L1 + l] = tw * (pref * t_exp<TR>(-gam * tr2));
                        }
                    }
                } else {
                    for (int i = 0; i < NL; ++i) Rrow[i] = TR(0);
                    const bool use_limit = !create_path && r < 1e-4;
                    if (use_limit) {
                        for (int k = 0; k < n; ++k)
#pragma unroll
                            for (int l = 0; l < L1; ++l) Rrow[k * L1 + l] = (TR)(w * b.poly_B0[k * L1 + l]);
                    } else {
                        const int Q = b.quad_n;
                        poly_levels(b, r, LMAX, [&](int q, int l, double P) {
                            for (int k = 0; k < n; ++k) Rrow[k * L1 + l] += (TR)(P * b.poly_gss[k * Q + q]);
                        });
                        double rl = 1.0;
#pragma unroll
                        for (int l = 0; l < L1; ++l) {
                            for (int k = 0; k < n; ++k) Rrow[k * L1 + l] = (TR)(w * ((double)Rrow[k * L1 + l] / rl));
                            rl *= r;
                        }
                    }
                }
            }
            __syncthreads();
            // ---- accumulate
            for (int e = 0; e < ntile; ++e) {
                const TR* Rr = Rsh + e * NL;
                const TY* Yr = Ysh + e * L2;
#pragma unroll
                for (int i = 0; i < NSLOT; ++i) {
                    if (lane + 64 * i < nslots) acc[i] += Rr[slot_r[i]] * (TR)Yr[slot_y[i]];
                }
            }
            __syncthreads();
            // ---- keep the remainder of the gathered list
            const int rem = cnt - ntile;
            int keep = 0;
            if (lane < rem) keep = esh[kTile + lane];
            __syncthreads();
            if (lane < rem) esh[lane] = keep;
            cnt = rem;
            __syncthreads();
        }
        // ---- finalise species s
        const size_t obase = ((size_t)c * S + s) * (size_t)nslots;
        const bool self_here = create_path && self_sp[c] == s;
        if (RBF == CUSOAP_RBF_GTO) {
#pragma unroll
            for (int i = 0; i < NSLOT; ++i)
                if (lane + 64 * i < nslots) accsh[lane + 64 * i] = acc[i];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < NSLOT; ++i) {
                const int t = lane + 64 * i;
                if (t < nslots) {
                    const int nn = t / L2, lm = slot_y[i], l = slot_r[i] - nn * L1;
                    const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
                    TR v = TR(0);
                    for (int k = 0; k < n; ++k) v += (TR)beta[k] * accsh[k * L2 + lm];
                    if (self_here && lm == 0) v += (TR)b.self_l0[nn];
                    out[obase + t] = v;
                }
            }
            __syncthreads();
        } else {
#pragma unroll
            for (int i = 0; i < NSLOT; ++i) {
                const int t = lane + 64 * i;
                if (t < nslots) {
                    TR v = acc[i];
                    if (self_here && slot_y[i] == 0) v += (TR)b.self_l0[t / L2];
                    out[obase + t] = v;
                }
            }
        }
    }
}

template <int LMAX, typename TY, typename TR, int RBF>
int launch_coeff(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                 const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                 const int64_t* offsets, const int32_t* nbr, int create_path, void* out, hipStream_t st) {
    using D = Dims<LMAX>;
    const int n = b.n_max;
    const size_t smem = sizeof(TR) * ((size_t)kTile * n * D::L1 + (size_t)n * D::L2) +
                        sizeof(TY) * (size_t)kTile * D::L2 + sizeof(int32_t) * 2 * kTile;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "coefficients: LDS tile too large");
    auto kern = coeff_kernel<LMAX, TY, TR, RBF>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("coefficients", st);
    kern<<<n_centers, 64, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, self_sp, offsets, nbr,
                                      create_path, (TR*)out);
    prof_end("coefficients", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <int LMAX>
int dispatch_prec(const cusoap_basis_t& b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                  const double* shifts, const double* centers, int n_centers, const int32_t* self_sp,
                  const int64_t* offsets, const int32_t* nbr, int create_path, void* out, hipStream_t st) {
#define CUSOAP_COEFF_ARGS b, pos, n_atoms, atom_sp, shifts, centers, n_centers, self_sp, offsets, nbr, create_path, out, st
    if (b.rbf == CUSOAP_RBF_GTO) {
        if (prec == CUSOAP_PREC_F64) return launch_coeff<LMAX, double, double, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
        if (prec == CUSOAP_PREC_F32) return launch_coeff<LMAX, float, float, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
        return launch_coeff<LMAX, float, double, CUSOAP_RBF_GTO>(CUSOAP_COEFF_ARGS);
    }
    if (prec == CUSOAP_PREC_F64) return launch_coeff<LMAX, double, double, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
    if (prec == CUSOAP_PREC_F32) return launch_coeff<LMAX, float, float, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
    return launch_coeff<LMAX, float, double, CUSOAP_RBF_POLY>(CUSOAP_COEFF_ARGS);
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
