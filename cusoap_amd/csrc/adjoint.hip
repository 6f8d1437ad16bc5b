// Adjoint of the descriptor with respect to the positions (vector-Jacobian
// product), without forming d desc / d x.  No reference counterpart: the
// reference only builds the dense Jacobian (soap_generator.py:2270-2626), which
// is infeasible beyond ~10^4 atoms.
//
//   L = sum_{c,f} G_desc[c][f] desc[c][f]
//   G[c][s][n][lm] = dL/dc:  power spectrum  pref_l sum_{s'n'} W_l[(s n),(s' n')] c[s'][n'][lm]
//                            (W from G_desc through the pair -> feature table, x2 on the diagonal)
//                            cc              G_desc permuted back to (y, z, x) and scaled
//   dL/dx_atom = sum over edges (c -> atom) of sum_{n,lm} G[c][s][n][lm] g[d][n][lm],
//   g the edge gradient of the closed-form Jacobian (fixed centers):
//   g[d][n][lm] = r_d Bt_nl Y_lm + Bw_nl dY_lm/dr_d.
//
// Kernel 1 forms G and stores M[c][s][lm][k]: for F64/F32 with the Loewdin
// matrix folded in, M = sum_n beta_l[n][k] G[s][n][lm], so kernel 2 contracts
// with the primitive radial functions directly; for MIXED and the radial table,
// M = G.  For the GTO / table edge kernels M also carries the harmonic
// normalisation K_lm.  With one descriptor gradient for all centers (the
// linear-readout case) W_l is built once and a grid-stride kernel forms G from
// it; otherwise one workgroup per center builds its own W_l.
// Kernel 2 (GTO; thread per edge, 256 threads per group of up to 8 centers whose
// edges form one contiguous range, so lanes stay busy across centers) walks the solid
// harmonics l-major, builds the radial vectors of degree l once per degree and
// contracts: H0 = sum_k rad0_k M[s][lm][k], H1 = sum_k rad1_k M[s][lm][k]; the
// gradient sum_lm [r H1 Y_lm + H0 grad Y_lm] is accumulated as r P + (X, Y, Z)
// from the shared factors of each (l, |m|) pair (solid_harmonics_adjoint), and
// the 3-vector is scattered with float64 atomics.  The polynomial basis uses a
// float64 tile kernel.
//
// Fixed-order force sums (every atom its own center): the neighbour lists are then
// symmetric, and rev[e] (edge_reverse_kernel, a binary search in the other atom's
// range, which is sorted by extended index) is the position of the reverse edge.  The
// edge pass stores each edge's contribution at its own index (coalesced), and
// vjp_gather_kernel sums, for atom a, the contributions of the edges c -> a, found as
// rev of the entries of a's own range, in float64 in a fixed order: the forces are the
// same bits on every run.  An edge whose reverse is missing (a pair within rounding of
// the cutoff, whose distance the two lists evaluate differently) adds its
// contribution with a float64 atomic instead.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include "common.hpp"

namespace cusoap {

// Buffers of the fixed-order force sums (all null: atomics).  Edge indices are those of
// offsets / nbr; e_max bounds them (the kernels also check offsets[n_centers] <= e_max).
struct VjpOrder {
    const int32_t* rev;   // (e_max) position of each edge's reverse edge, or -1 (atomic)
    void* slots;          // (e_max, 3) each edge's contribution, float or double as the edge pass
    double* att;          // (n_centers, 3) sum of each attached center's edge contributions
    int64_t e_max;
};

namespace {

__device__ __forceinline__ bool vjp_ordered(const VjpOrder& o, const int64_t* offsets, int n_centers) {
    return o.rev != nullptr && offsets[n_centers] <= o.e_max;
}

__device__ __forceinline__ int lm_l2(int lm) {
    int l = (int)sqrtf((float)lm);
    while (l * l > lm) --l;
    while ((l + 1) * (l + 1) <= lm) ++l;
    return l;
}

// K_lm of the real solid harmonic lm (folded into M for the GTO edge kernels)
__device__ __forceinline__ double sh_norm_lm(int lm) {
    const int l = lm_l2(lm), m = lm - l * l - l;
    return kSHNorm[l][m < 0 ? -m : m];
}

// our lm -> DScribe cc slot holding it (inverse of the (z, x, y) permutation)
__device__ __forceinline__ int cc_slot_of(int lm) {
    if (lm == 1) return 3;
    if (lm == 2) return 1;
    if (lm == 3) return 2;
    return lm;
}

// M[c][s][lm][k] (TM): dL/dc, with beta folded in when fold_beta.
// Power spectrum: the workgroup first forms W_l[a][b] = (2 if a == b else 1)
// g[feature(l, a, b)] in LDS from the symmetric feature table (coalesced reads),
// then G[a][lm] = pref_l sum_b W_l[a][b] c[b][lm] from LDS only.
template <typename TC, typename TG, typename TM>
__global__ __launch_bounds__(256) void coeff_adjoint_kernel(cusoap_basis_t b, int mode, const TC* __restrict__ coeffs,
                                                            const int32_t* __restrict__ pair_feature,
                                                            const TG* __restrict__ gd, int gd_broadcast, int n_feat,
                                                            int fold_beta, int fold_k, TM* __restrict__ M) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int S = b.n_species, n = b.n_max, L1 = b.l_max + 1, L2 = L1 * L1;
    const int per = S * n * L2, SN = S * n;
    double* cs = (double*)smem;       // [per]   c[s][n][lm]
    double* G = cs + per;              // [per]   G[s][n][lm]
    double* g = G + per;               // [n_feat]
    const int WS = SN + 1;             // padded W row: lanes on different rows hit different banks
    double* W = g + n_feat;            // [L1][SN][WS] (power spectrum)
    double* bs = W + (mode == CUSOAP_MODE_PS ? L1 * SN * WS : 0);   // [L1][n][n] beta (fold)
    const int c = blockIdx.x, tid = threadIdx.x;
    const TC* src = coeffs + (size_t)c * per;
    const TG* gsrc = gd + (gd_broadcast ? 0 : (size_t)c * n_feat);
    for (int i = tid; i < per; i += blockDim.x) cs[i] = (double)src[i];
    for (int i = tid; i < n_feat; i += blockDim.x) g[i] = (double)gsrc[i];
    if (fold_beta)
        for (int i = tid; i < L1 * n * n; i += blockDim.x) bs[i] = b.gto_beta[i];
    __syncthreads();
    if (mode == CUSOAP_MODE_PS) {
        __shared__ double prefl[kMaxL + 1];
        if (tid < L1) prefl[tid] = sqrt(8.0 * M_PI * M_PI / (2 * tid + 1));
        __syncthreads();
        const int nw = L1 * SN * SN;
        for (int i = tid; i < nw; i += blockDim.x) {
            const int f = pair_feature[i];
            const int l = i / (SN * SN), ab = i - l * SN * SN;
            const int a = ab / SN, bb = ab - a * SN;
            W[((size_t)l * SN + a) * WS + bb] = f >= 0 ? (bb == a ? 2.0 : 1.0) * prefl[l] * g[f] : 0.0;
        }
        __syncthreads();
    }
    for (int t = tid; t < per; t += blockDim.x) {
        const int a = t / L2, lm = t - a * L2, l = lm_l2(lm);
        double v = 0.0;
        if (mode == CUSOAP_MODE_PS) {
            const double* wr = W + ((size_t)l * SN + a) * WS;
            for (int bb = 0; bb < SN; ++bb) v += wr[bb] * cs[bb * L2 + lm];
        } else {
            const int slot = cc_slot_of(lm);
            v = g[a * L2 + slot] * (slot >= 4 ? 0.17958712212516656 : 1.0);
        }
        G[t] = v;
    }
    __syncthreads();
    TM* Mc = M + (size_t)c * per;
    for (int t = tid; t < per; t += blockDim.x) {
        // output order [s][lm][k]
        const int s = t / (L2 * n), rem = t - s * L2 * n;
        const int lm = rem / n, k = rem - lm * n;
        double v;
        if (fold_beta) {
            const int l = lm_l2(lm);
            const double* br = bs + (size_t)l * n * n + k;
            v = 0.0;
            for (int nn = 0; nn < n; ++nn) v += br[nn * n] * G[(s * n + nn) * L2 + lm];
        } else {
            v = G[(s * n + k) * L2 + lm];
        }
        if (fold_k) v *= sh_norm_lm(lm);
        Mc[t] = (TM)v;
    }
}

// LDS layout of W_l[a][b] in the shared-gradient kernel: odd row and degree
// strides, so that the lanes of a wavefront (different l and a) spread over banks
__host__ __device__ inline int psw_row(int SN) { return SN + 1 - (SN & 1); }
__host__ __device__ inline int psw_deg(int SN) { return (SN * psw_row(SN)) | 1; }
__host__ __device__ inline int psw_size(int L1, int SN) { return L1 * psw_deg(SN) + 3 * psw_row(SN); }

// W_l[a][b] = (2 if a == b else 1) pref_l g[feature(l, a, b)] once for all centers
// (descriptor gradient shared by every center, the linear-readout case), in the
// padded layout above with three zero rows past the end (row-quad overrun).
template <typename TG, typename TW>
__global__ void ps_weights_kernel(int L1, int SN, const int32_t* __restrict__ pair_feature, const TG* __restrict__ gd,
                                  TW* __restrict__ W) {
    const int RS = psw_row(SN), LS = psw_deg(SN), nw = psw_size(L1, SN);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
        const int l = i / LS, rem = i - l * LS, a = rem / RS, bb = rem - a * RS;
        TW v = TW(0);
        if (l < L1 && a < SN && bb < SN) {
            const int f = pair_feature[((size_t)l * SN + a) * SN + bb];
            const double prefl = sqrt(8.0 * M_PI * M_PI / (2 * l + 1));
            if (f >= 0) v = (TW)((bb == a ? 2.0 : 1.0) * prefl * (double)gd[f]);
        }
        W[i] = v;
    }
}

// Power-spectrum adjoint for a shared descriptor gradient.  A grid-stride loop
// over centers; W (from ps_weights_kernel) is staged in LDS once per workgroup,
// and the next center's coefficients are loaded into registers while the
// current one is processed, so the loop is not one global-load latency per center.
// Thread (lm, quad of rows a) forms G[a][lm] = sum_b W_l[a][b] c[b][lm] for four
// a at once (one c read per four products), b ascending as in the per-center
// kernel; then M is written in [s][lm][k] order (beta folded in when asked).
// TW: arithmetic type (float64 when beta is folded or M is float64).
constexpr int kPsPre = 8;        // prefetched values per thread (rows up to 2048 values)
template <typename TC, typename TW, typename TM>
__global__ __launch_bounds__(256) void ps_adjoint_shared_kernel(int S, int n, int L1, const TC* __restrict__ coeffs,
                                                                const TW* __restrict__ Wg,
                                                                const double* __restrict__ beta, int fold_beta,
                                                                int fold_k, int n_centers, TM* __restrict__ M) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int SN = S * n, L2 = L1 * L1, per = SN * L2;
    const int RS = psw_row(SN), LS = psw_deg(SN), nw = psw_size(L1, SN);
    TW* W = (TW*)smem;                 // padded [L1][SN][SN] + 3 zero rows
    TW* cs = W + nw;                   // [SN][L2]
    TW* G = cs + per;                  // [SN][L2]
    TW* bs = G + per;                  // [L1][n][n] (fold)
    const int tid = threadIdx.x;
    for (int i = tid; i < nw; i += 256) W[i] = Wg[i];
    if (fold_beta)
        for (int i = tid; i < L1 * n * n; i += 256) bs[i] = (TW)beta[i];
    const int nq = (SN + 3) / 4;
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
            if (i < per) cs[i] = (TW)pre[j];
        }
        __syncthreads();
        fetch(c + gridDim.x);          // in flight while this center is processed
        for (int t = tid; t < L2 * nq; t += 256) {
            const int aq = t / L2, lm = t - aq * L2, l = lm_l2(lm), a0 = 4 * aq;
            const TW* w0 = W + (size_t)l * LS + a0 * RS;
            TW acc[4] = {TW(0), TW(0), TW(0), TW(0)};
            for (int bb = 0; bb < SN; ++bb) {
                const TW cv = cs[bb * L2 + lm];
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j] += w0[j * RS + bb] * cv;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (a0 + j < SN) G[(a0 + j) * L2 + lm] = acc[j];
        }
        __syncthreads();
        TM* Mc = M + (size_t)c * per;
        for (int t = tid; t < per; t += 256) {
            const int s = t / (L2 * n), rem = t - s * L2 * n;
            const int lm = rem / n, k = rem - lm * n;
            TW v;
            if (fold_beta) {
                const TW* br = bs + (size_t)lm_l2(lm) * n * n + k;
                v = TW(0);
                for (int nn = 0; nn < n; ++nn) v += br[nn * n] * G[(s * n + nn) * L2 + lm];
            } else {
                v = G[(s * n + k) * L2 + lm];
            }
            if (fold_k) v *= (TW)sh_norm_lm(lm);
            Mc[t] = (TM)v;
        }
        __syncthreads();
    }
}

// Power-spectrum adjoint on the matrix cores, float32 coefficients and M with no
// beta folded in (the mixed and radial-table forms).  Per center and degree l,
// G^T[m][a] = sum_b c[b][l^2 + m] W_l[a][b] is a (2l+1) x SN x SN product:
// v_mfma_f32_16x16x4_f32 in K steps of 4 over b (float32 products and float32 sums,
// the arithmetic of the vector kernel in another order), rows m in tiles of 16.
// Each wavefront takes one center at a time: its coefficient block is copied to the
// wavefront's LDS slice with coalesced reads, the products read A (coefficients)
// and B = W_l^T (staged once per workgroup) from LDS, the results are scattered
// into the [s][lm][k] order of M in a second LDS slice (with K_lm when fold_k), and
// the block is written out with coalesced stores.  Lane layout of the instruction:
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i / 4),
// element i % 4.
typedef float mf32x4 __attribute__((ext_vector_type(4)));
__host__ __device__ inline int ps_mfma_slice(int per) { return (per + 3) & ~3; }
constexpr int kPsMfmaPre = 32;           // prefetched values per lane (blocks up to 2048 values)
constexpr int kPsMfmaMaxNT = 4;          // output tiles of 16 (species x n_max up to 64)
__global__ __launch_bounds__(256) void ps_adjoint_mfma_kernel(int S, int n, int L1, const float* __restrict__ coeffs,
                                                              const float* __restrict__ Wg, int fold_k, int n_centers,
                                                              float* __restrict__ M) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int SN = S * n, L2 = L1 * L1, per = SN * L2, slice = ps_mfma_slice(per);
    const int RS = psw_row(SN), LS = psw_deg(SN), nw = psw_size(L1, SN);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    float* W = (float*)smem;                       // padded [L1][SN][SN] + 3 zero rows
    float* knorm = W + ((nw + 3) & ~3);            // [L2] K_lm (or 1)
    float* cin = knorm + ((L2 + 3) & ~3) + wv * 2 * slice;   // this wavefront's coefficient block
    float* cout = cin + slice;                     // and its M block
    for (int i = tid; i < nw; i += 256) W[i] = Wg[i];
    for (int i = tid; i < L2; i += 256) knorm[i] = fold_k ? (float)sh_norm_lm(i) : 1.f;
    __syncthreads();
    const int NT = (SN + 15) >> 4, KS = (SN + 3) >> 2;
    const int i16 = lane & 15, kk = lane >> 4;
    // this lane's output a = 16 nt + i16 as (species, k) offsets into the M block
    int aoff[kPsMfmaMaxNT];
#pragma unroll
    for (int nt = 0; nt < kPsMfmaMaxNT; ++nt) {
        const int a = 16 * nt + i16, sa = a / n;
        aoff[nt] = a < SN ? sa * L2 * n + (a - sa * n) : -1;
    }
    const int stride = gridDim.x * 4;
    // the next center's block is loaded into registers while this one is processed
    float pre[kPsMfmaPre];
    auto fetch = [&](int c) {
        if (c >= n_centers) return;
        const float* src = coeffs + (size_t)c * per;
#pragma unroll
        for (int j = 0; j < kPsMfmaPre; ++j) {
            const int i = lane + 64 * j;
            if (i < per) pre[j] = src[i];
        }
    };
    fetch(blockIdx.x * 4 + wv);
    for (int c = blockIdx.x * 4 + wv; c < n_centers; c += stride) {
#pragma unroll
        for (int j = 0; j < kPsMfmaPre; ++j) {
            const int i = lane + 64 * j;
            if (i < per) cin[i] = pre[j];
        }
        fetch(c + stride);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int l = 0; l < L1; ++l) {
            const int d = 2 * l + 1, l2 = l * l;
#pragma unroll
            for (int nt = 0; nt < kPsMfmaMaxNT; ++nt) {
                if (nt >= NT) break;
                const int a = 16 * nt + i16;       // this lane's B column / output a
                const bool a_ok = a < SN;
                const float* Wa = W + l * LS + a * RS;
                const int ao = aoff[nt];
                for (int m0 = 0; m0 < d; m0 += 16) {
                    const int mr = m0 + i16;       // this lane's A row
                    mf32x4 acc = {0.f, 0.f, 0.f, 0.f};
                    for (int ks = 0; ks < KS; ++ks) {
                        const int bb = 4 * ks + kk;
                        const bool b_ok = bb < SN;
                        const float av = (mr < d && b_ok) ? cin[bb * L2 + l2 + mr] : 0.f;
                        const float bv = (a_ok && b_ok) ? Wa[bb] : 0.f;
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
                    }
                    if (a_ok) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int m = m0 + 4 * kk + v;
                            if (m < d) cout[ao + (l2 + m) * n] = acc[v] * knorm[l2 + m];
                        }
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float* dst = M + (size_t)c * per;
        for (int i = lane; i < per; i += 64) dst[i] = cout[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// CUSOAP_PS_ADJOINT=vector selects the vector kernel instead (read at each call)
inline bool ps_mfma_enabled() {
    const char* v = std::getenv("CUSOAP_PS_ADJOINT");
    return !(v && std::strcmp(v, "vector") == 0);
}

template <typename T>
__device__ __forceinline__ T t_exp2(T v);
template <>
__device__ __forceinline__ float t_exp2<float>(float v) { return expf(v); }
template <>
__device__ __forceinline__ double t_exp2<double>(double v) { return exp(v); }

// GTO adjoint, thread per edge.  TM: type of M and of the contraction.
// RAD: kRadPrim  M has beta folded in; radial vectors are the primitives.
//      kRadMixed M = G (no beta), radial B/Bt per edge in float64 then rounded to TM.
//      kRadTable M = G, B and Bt = (1/r) dB/dr by cubic Hermite interpolation of
//                basis.rad_table (the derivative of the interpolant, so forward and
//                adjoint describe the same function).
// The radial vectors of the current degree live in LDS columns [k][thread], or,
// with a compile-time n_max (NN > 0), in registers, so that each lm needs only
// the 16-byte reads of its M row.
constexpr int kRadPrim = 0, kRadMixed = 1, kRadTable = 2;
// centers per workgroup of the edge kernel (at most): the ALU-bound forms keep all
// lanes busy with groups of 8; the register-table form is bound by the latency of
// its table reads and measured fastest with one center per workgroup
constexpr int kVjpGroup = 8;
// register-table form: up to 16 neighbouring centers per workgroup, their force contributions
// summed per distinct neighbour atom in an LDS hash table (float sums, linear probing)
// before one float64 atomic per atom and component: the neighbourhoods of consecutive
// centers overlap, so this cuts the global atomics (executed at the memory side).
// 8 centers against 4: water box edge pass 2.12 -> 1.96 ms, 1M cluster 12.7 -> 11.1 ms;
// 16 (as many as 64 KB of coefficient rows allow): 1M cluster 11.1 -> 10.5 ms, water box
// unchanged (its rows allow 16 too)
constexpr int kVjpHashGroup = 16;
constexpr int kVjpMaxGroup = kVjpGroup > kVjpHashGroup ? kVjpGroup : kVjpHashGroup;
constexpr int kVjpHashSlots = 1024;
constexpr int kVjpHashProbe = 32;      // probes before an edge falls back to global atomics
template <int NN>
constexpr int vjp_group() { return NN > 0 ? kVjpHashGroup : kVjpGroup; }
// the register-table form, which prefetches the next degree's table rows, was
// measured fastest at 3 wavefronts per SIMD (2.35 ms against 2.47 at 4 and 2.36 unbounded)
template <int NN>
constexpr int vjp_min_waves() { return NN > 0 ? 3 : 1; }

// M row of one lm into registers (16-byte reads when the row allows)
template <typename T, int N>
__device__ __forceinline__ void load_row(const T* p, T (&v)[N]) {
    if constexpr (sizeof(T) == 4 && N % 4 == 0) {
#pragma unroll
        for (int k = 0; k < N; k += 4) {
            const float4 q = *(const float4*)(p + k);
            v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = p[k];
    }
}

template <int LMAX, typename TM, int RAD, int NN>
__global__ __launch_bounds__(256, vjp_min_waves<NN>()) void vjp_gto_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                      const int32_t* __restrict__ atom_sp,
                                                      const double* __restrict__ shifts,
                                                      const double* __restrict__ centers,
                                                      const int32_t* __restrict__ center_attach,
                                                      const int32_t* __restrict__ center_species,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ nbr, const TM* __restrict__ M,
                                                      const int32_t* __restrict__ order, int n_centers, int ngrp,
                                                      int per_xcd, double* __restrict__ grad, VjpOrder ord) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    constexpr int NT = 256;
    constexpr bool MIXED = RAD == kRadMixed;
    constexpr bool FIXN = NN > 0;
    constexpr int NR = FIXN ? NN : 1;
    // node rows in registers: (value, slope) pairs of two n per float4, or of one n per
    // float2 when n is odd (rows of 2n floats are then 8-byte aligned only)
    constexpr bool ODDN = (NN & 1) != 0;
    using RowV = typename std::conditional<ODDN, float2, float4>::type;
    constexpr int NQ = FIXN && RAD == kRadTable ? (ODDN ? NN : NN / 2) : 1;
    const int n = FIXN ? NN : b.n_max, S = b.n_species;
    extern __shared__ __align__(16) unsigned char smem[];
    double* qd = (double*)smem;                             // [n][NT]   (MIXED, runtime n) primitives
    TM* rad0 = (TM*)(qd + (MIXED && !FIXN ? n * NT : 0));   // [n][NT]   (runtime n)
    TM* rad1 = rad0 + (FIXN ? 0 : n * NT);                  // [n][NT]   (runtime n)
    TM* Ms = rad1 + (FIXN ? 0 : n * NT);                    // [ngrp][S][L2][n]
    __shared__ int gid[kVjpMaxGroup];                       // the group's centers (order[c0 + i])
    __shared__ int64_t gbeg[kVjpMaxGroup];                  // first edge of each
    __shared__ int64_t goff[kVjpMaxGroup + 1];              // their edges' offsets in the group's flat range
    __shared__ double gcen[kVjpMaxGroup][3];
    __shared__ int gslot[kVjpMaxGroup];                        // table slot of each center
    __shared__ double red[NT / 64][kVjpMaxGroup][3];           // center-attach sums per wavefront
    constexpr bool HASH = FIXN && RAD == kRadTable;
    __shared__ int hkey[HASH ? kVjpHashSlots : 1];          // neighbour atom of each slot or -1
    __shared__ float hval[HASH ? kVjpHashSlots : 1][3];     // its summed contributions
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool ordered = vjp_ordered(ord, offsets, n_centers);
    TM* const slots = (TM*)ord.slots;
    if (HASH && !ordered)
        for (int i = tid; i < kVjpHashSlots; i += NT) {
            hkey[i] = -1;
            hval[i][0] = hval[i][1] = hval[i][2] = 0.0f;
        }
    // the group: centers c0 .. c0 + ng - 1 of the order (the identity without one);
    // consecutive workgroups go to different XCDs, each XCD taking a contiguous range of
    // groups, whose neighbourhoods then share its L2
    const int grp = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    const int c0 = grp * ngrp;
    if (c0 >= n_centers) return;
    const int ng = n_centers - c0 < ngrp ? n_centers - c0 : ngrp;
    const int per = S * n * L2;
    if (tid < ng) {
        const int c = order ? order[c0 + tid] : c0 + tid;
        gid[tid] = c;
        gbeg[tid] = offsets[c];
        goff[tid + 1] = offsets[c + 1] - offsets[c];            // counts; the prefix follows
        for (int d = 0; d < 3; ++d) {
            gcen[tid][d] = centers[3 * (size_t)c + d];
            for (int w = 0; w < NT / 64; ++w) red[w][tid][d] = 0.0;
        }
        gslot[tid] = (RAD == kRadTable && center_species[c] >= 0) ? center_species[c] : S;
    }
    __syncthreads();
    if (tid == 0) {
        goff[0] = 0;
        for (int i = 0; i < ng; ++i) goff[i + 1] += goff[i];
    }
    {
        // the group's coefficient rows, center by center: 16-byte loads (rows whose
        // length is a multiple of 16 bytes), four in flight per thread
        constexpr int V = 16 / sizeof(TM);
        using Vec = typename std::conditional<sizeof(TM) == 4, float4, double2>::type;
        if (per % V == 0) {
            const int pv = per / V, nv = ng * pv;
            Vec* d4 = (Vec*)Ms;
            auto src = [&](int i) {
                const int k = i / pv;
                return ((const Vec*)(M + (size_t)gid[k] * per))[i - k * pv];
            };
            int i = tid;
            for (; i + 3 * NT < nv; i += 4 * NT) {
                const Vec a0 = src(i), a1 = src(i + NT), a2 = src(i + 2 * NT), a3 = src(i + 3 * NT);
                d4[i] = a0;
                d4[i + NT] = a1;
                d4[i + 2 * NT] = a2;
                d4[i + 3 * NT] = a3;
            }
            for (; i < nv; i += NT) d4[i] = src(i);
        } else {
            for (int j = tid; j < ng * per; j += NT) {
                const int k = j / per;
                Ms[j] = M[(size_t)gid[k] * per + (j - k * per)];
            }
        }
    }
    __syncthreads();
    const int64_t E1 = goff[ng];
    // table interval width and its inverse (once; the edges multiply)
    const double hd = RAD == kRadTable ? b.rad_rmax / b.rad_intervals : 0.0;
    const double ihd = RAD == kRadTable ? b.rad_intervals / b.rad_rmax : 0.0;
    TM* r0c = rad0 + tid;
    TM* r1c = rad1 + tid;
    double* qc = qd + tid;
    // the group's edges as one flat range (lanes stay busy across center boundaries)
    for (int64_t base = 0; base < E1; base += NT) {
        const int64_t f = base + tid;
        int ci = -1;
        TM gx = TM(0), gy = TM(0), gz = TM(0);
        if (f < E1) {
            ci = 0;
            while (ci + 1 < ng && goff[ci + 1] <= f) ++ci;
            const int64_t e = gbeg[ci] + (f - goff[ci]);
            const double cx = gcen[ci][0], cy = gcen[ci][1], cz = gcen[ci][2];
            const int cslot = gslot[ci];
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
            const double ir = r > 0.0 ? 1.0 / r : 0.0;            // the edge's one division
            const double dwr = dw * ir;
            const TM* Mrow = Ms + (size_t)ci * per + (size_t)atom_sp[a] * L2 * n;
            const TM trx = (TM)rx, tryy = (TM)ry, trz = (TM)rz;
            // table interpolation weights: value (v*) and (1/r) d/dr (c*)
            const float* tb = nullptr;
            TM v00 = TM(0), v10 = TM(0), v01 = TM(0), v11 = TM(0), c00 = TM(0), c10 = TM(0), c01 = TM(0), c11 = TM(0);
            if (RAD == kRadTable) {
                const int G = b.rad_intervals;
                const double xr = r * ihd;
                int i = (int)xr;
                i = i < G - 1 ? i : G - 1;
                const double t = xr - i, t2 = t * t, t3 = t2 * t;
                v00 = (TM)(2.0 * t3 - 3.0 * t2 + 1.0);
                v01 = (TM)(3.0 * t2 - 2.0 * t3);
                v10 = (TM)(hd * (t3 - 2.0 * t2 + t));
                v11 = (TM)(hd * (t3 - t2));
                if (i == 0) {
                    // the node slope at r = 0 is zero, so P'(r)/r is a polynomial on the first interval
                    c00 = (TM)(6.0 * (t - 1.0) * (ihd * ihd));
                    c01 = -c00;
                    c11 = (TM)((3.0 * t - 2.0) * ihd);
                } else {
                    c00 = (TM)((6.0 * t2 - 6.0 * t) * ihd * ir);
                    c10 = (TM)((3.0 * t2 - 4.0 * t + 1.0) * ir);
                    c01 = -c00;
                    c11 = (TM)((3.0 * t2 - 2.0 * t) * ir);
                }
                tb = b.rad_table + ((size_t)(cslot * S + atom_sp[a]) * (G + 1) + i) * (L1 * 2 * n);
                // fold the weighting in: rad0 = w P(r), rad1 = w P'(r)/r + (w'/r) P(r)
                const TM tw = (TM)w, tdwr = (TM)dwr;
                c00 = tw * c00 + tdwr * v00;
                c10 = tw * c10 + tdwr * v10;
                c01 = tw * c01 + tdwr * v01;
                c11 = tw * c11 + tdwr * v11;
                v00 *= tw;
                v10 *= tw;
                v01 *= tw;
                v11 *= tw;
            }
            TM rr0[NR], rr1[NR];                                 // radial vectors of degree l (FIXN)
            // register table form: node rows of the next degree
            RowV nx0[NQ], nx1[NQ];
            auto load_rows = [&](int l) {
                if constexpr (FIXN && RAD == kRadTable) {
                    const RowV* a0 = (const RowV*)(tb + l * 2 * NN);
                    const RowV* a1 = (const RowV*)(tb + (L1 + l) * 2 * NN);
#pragma unroll
                    for (int k = 0; k < NQ; ++k) {
                        nx0[k] = a0[k];
                        nx1[k] = a1[k];
                    }
                }
            };
            load_rows(0);
            auto on_l = [&](int l) {
                if constexpr (FIXN) {
                    const TM tw = (TM)w, tdwr = (TM)dwr;
                    if (RAD == kRadTable) {
                        // this degree's rows were loaded one degree ahead; issue the next degree's
                        // loads before using them, so their latency overlaps this degree's work
                        RowV c0[NQ], c1[NQ];
#pragma unroll
                        for (int k = 0; k < NQ; ++k) {
                            c0[k] = nx0[k];
                            c1[k] = nx1[k];
                        }
                        if (l + 1 < L1) load_rows(l + 1);
#pragma unroll
                        for (int k = 0; k < NN; ++k) {
                            TM b0, d0, b1, d1;
                            if constexpr (ODDN) {
                                b0 = (TM)c0[k].x; d0 = (TM)c0[k].y; b1 = (TM)c1[k].x; d1 = (TM)c1[k].y;
                            } else {
                                const float4 q0 = c0[k >> 1], q1 = c1[k >> 1];
                                b0 = (TM)((k & 1) ? q0.z : q0.x); d0 = (TM)((k & 1) ? q0.w : q0.y);
                                b1 = (TM)((k & 1) ? q1.z : q1.x); d1 = (TM)((k & 1) ? q1.w : q1.y);
                            }
                            rr0[k] = v00 * b0 + v10 * d0 + v01 * b1 + v11 * d1;
                            rr1[k] = c00 * b0 + c10 * d0 + c01 * b1 + c11 * d1;
                        }
                    } else if (MIXED) {
                        double qv[NN];
    #pragma unroll
                        for (int k = 0; k < NN; ++k) qv[k] = b.gto_pref[l * NN + k] * exp(-b.gto_gamma[l * NN + k] * r2);
    #pragma unroll
                        for (int nn = 0; nn < NN; ++nn) {
                            const double* beta = b.gto_beta + ((size_t)l * NN + nn) * NN;
                            double B = 0.0, Bh = 0.0;
    #pragma unroll
                            for (int k = 0; k < NN; ++k) {
                                B += beta[k] * qv[k];
                                Bh += beta[k] * (-2.0 * b.gto_gamma[l * NN + k]) * qv[k];
                            }
                            rr0[nn] = (TM)(w * B);
                            rr1[nn] = (TM)(w * Bh + dwr * B);
                        }
                    } else {
    #pragma unroll
                        for (int k = 0; k < NN; ++k) {
                            const TM gam = (TM)b.gto_gamma[l * NN + k];
                            const TM q = (TM)b.gto_pref[l * NN + k] * t_exp2<TM>(-gam * (TM)r2);
                            rr0[k] = tw * q;
                            rr1[k] = (tw * (TM(-2) * gam) + tdwr) * q;
                        }
                    }
                    return;
                }
                if (RAD == kRadTable) {
                    const float* p0 = tb + l * 2 * n;
                    const float* p1 = p0 + L1 * 2 * n;
                    auto emit = [&](int nn, float b0, float d0, float b1, float d1) {
                        r0c[nn * NT] = v00 * (TM)b0 + v10 * (TM)d0 + v01 * (TM)b1 + v11 * (TM)d1;
                        r1c[nn * NT] = c00 * (TM)b0 + c10 * (TM)d0 + c01 * (TM)b1 + c11 * (TM)d1;
                    };
                    if ((n & 1) == 0) {
                        // node rows start on 16-byte boundaries: two 16-byte loads per 2 functions
                        for (int k2 = 0; k2 < n; k2 += 2) {
                            const float4 a = *(const float4*)(p0 + 2 * k2), c = *(const float4*)(p1 + 2 * k2);
                            emit(k2, a.x, a.y, c.x, c.y);
                            emit(k2 + 1, a.z, a.w, c.z, c.w);
                        }
                    } else {
                        for (int nn = 0; nn < n; ++nn) emit(nn, p0[2 * nn], p0[2 * nn + 1], p1[2 * nn], p1[2 * nn + 1]);
                    }
                } else if (MIXED) {
                    for (int k = 0; k < n; ++k) qc[k * NT] = b.gto_pref[l * n + k] * exp(-b.gto_gamma[l * n + k] * r2);
                    for (int nn = 0; nn < n; ++nn) {
                        const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
                        double B = 0.0, Bh = 0.0;
                        for (int k = 0; k < n; ++k) {
                            const double q = qc[k * NT];
                            B += beta[k] * q;
                            Bh += beta[k] * (-2.0 * b.gto_gamma[l * n + k]) * q;
                        }
                        r0c[nn * NT] = (TM)(w * B);
                        r1c[nn * NT] = (TM)(w * Bh + dwr * B);
                    }
                } else {
                    for (int k = 0; k < n; ++k) {
                        const TM gam = (TM)b.gto_gamma[l * n + k];
                        const TM q = (TM)b.gto_pref[l * n + k] * t_exp2<TM>(-gam * (TM)r2);
                        r0c[k * NT] = (TM)w * q;
                        r1c[k * NT] = ((TM)w * (TM(-2) * gam) + (TM)dwr) * q;
                    }
                }
            };
            // the two radial contractions of lm with M (which carries K_lm)
            auto hf = [&](int lm, TM& h0, TM& h1) {
                const TM* mv = Mrow + lm * n;
                h0 = TM(0);
                h1 = TM(0);
                if constexpr (FIXN) {
                    TM mk[NR];
                    load_row<TM, NR>(mv, mk);
    #pragma unroll
                    for (int k = 0; k < NR; ++k) {
                        h0 += rr0[k] * mk[k];
                        h1 += rr1[k] * mk[k];
                    }
                } else {
                    for (int k = 0; k < n; ++k) {
                        const TM m = mv[k];
                        h0 += r0c[k * NT] * m;
                        h1 += r1c[k * NT] * m;
                    }
                }
            };
            TM P = TM(0), X = TM(0), Y = TM(0), Z = TM(0);
            solid_harmonics_adjoint<LMAX, TM>(trx, tryy, trz, on_l, hf, P, X, Y, Z);
            gx = trx * P + X;
            gy = tryy * P + Y;
            gz = trz * P + Z;
            bool done = false;
            if (ordered) {
                // gathered through the reverse edge; without one, by atomic below
                slots[3 * e] = gx;
                slots[3 * e + 1] = gy;
                slots[3 * e + 2] = gz;
                done = ord.rev[e] >= 0;
            } else if (HASH) {
                unsigned h = ((unsigned)a * 2654435761u) >> 22;      // 10-bit multiplicative hash
                for (int pr = 0; pr < kVjpHashProbe; ++pr, h = (h + 1) & (kVjpHashSlots - 1)) {
                    const int k = atomicCAS(&hkey[h], -1, a);
                    if (k == -1 || k == a) {
                        atomicAdd(&hval[h][0], (float)gx);
                        atomicAdd(&hval[h][1], (float)gy);
                        atomicAdd(&hval[h][2], (float)gz);
                        done = true;
                        break;
                    }
                }
            }
            if (!done) {
                atomicAdd(&grad[3 * (size_t)a], (double)gx);
                atomicAdd(&grad[3 * (size_t)a + 1], (double)gy);
                atomicAdd(&grad[3 * (size_t)a + 2], (double)gz);
            }
        }
        if (center_attach && __builtin_amdgcn_readfirstlane(ci) >= 0) {
            // the wavefront's edges are contiguous, so they span few centers: one masked
            // wavefront sum per center present, then one LDS add per wavefront and center
            for (int cc = __builtin_amdgcn_readfirstlane(ci); __any(ci >= cc); ++cc) {
                double v[3] = {ci == cc ? (double)gx : 0.0, ci == cc ? (double)gy : 0.0, ci == cc ? (double)gz : 0.0};
#pragma unroll
                for (int o = 32; o > 0; o >>= 1)
#pragma unroll
                    for (int d = 0; d < 3; ++d) v[d] += __shfl_xor(v[d], o, 64);
                if (lane == 0)
                    for (int d = 0; d < 3; ++d) red[wv][cc][d] += v[d];   // this wavefront's own sums
            }
        }
    }
    if (HASH || center_attach) __syncthreads();
    if (HASH && !ordered)
        for (int i = tid; i < kVjpHashSlots; i += NT) {
            const int a = hkey[i];
            if (a >= 0)
                for (int d = 0; d < 3; ++d) atomicAdd(&grad[3 * (size_t)a + d], (double)hval[i][d]);
        }
    if (center_attach && tid < ng) {
        const int c = gid[tid], ca = center_attach[c];
        if (ca >= 0) {
            double sum[3];
            for (int d = 0; d < 3; ++d) sum[d] = ((red[0][tid][d] + red[1][tid][d]) + red[2][tid][d]) + red[3][tid][d];
            if (ordered && ca == c)
                for (int d = 0; d < 3; ++d) ord.att[3 * (size_t)c + d] = sum[d];
            else
                for (int d = 0; d < 3; ++d) atomicAdd(&grad[3 * (size_t)ca + d], -sum[d]);
        }
    }
}

// polynomial basis, float64: one wavefront per center, tiles of 64 edges, lane = edge
template <int LMAX>
__global__ __launch_bounds__(64) void vjp_poly_kernel(cusoap_basis_t b, const double* __restrict__ pos, int n_atoms,
                                                      const int32_t* __restrict__ atom_sp,
                                                      const double* __restrict__ shifts,
                                                      const double* __restrict__ centers,
                                                      const int32_t* __restrict__ center_attach,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ nbr, const double* __restrict__ M,
                                                      const int32_t* __restrict__ order, int n_centers,
                                                      double* __restrict__ grad, VjpOrder ord) {
    constexpr int L1 = LMAX + 1, L2 = L1 * L1;
    const int n = b.n_max, S = b.n_species, NL = n * L1;
    const bool ordered = vjp_ordered(ord, offsets, n_centers);
    double* const slots = (double*)ord.slots;
    extern __shared__ __align__(16) unsigned char smem[];
    double* Ms = (double*)smem;               // [S][L2][n]
    double* R0 = Ms + S * n * L2;             // [64][NL]
    double* R1 = R0 + 64 * NL;                // [64][NL]
    const int c = order ? order[blockIdx.x] : (int)blockIdx.x, lane = threadIdx.x;
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
            int img, a;
            split_ext(ext, n_atoms, 1.0 / (double)n_atoms, img, a);
            const double rx = (pos[3 * (size_t)a] + shifts[3 * img]) - cx;
            const double ry = (pos[3 * (size_t)a + 1] + shifts[3 * img + 1]) - cy;
            const double rz = (pos[3 * (size_t)a + 2] + shifts[3 * img + 2]) - cz;
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            const double r = sqrt(r2);
            double w = 1.0, dw = 0.0;
            if (b.weighting != CUSOAP_W_NONE) weighting(b, r, w, dw);
            const double dwr = r > 0.0 ? dw / r : 0.0;
            const int s = atom_sp[a];
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
            const double* Mrow = Ms + (size_t)s * L2 * n;
            double gx = 0.0, gy = 0.0, gz = 0.0;
            solid_harmonics<LMAX, true, double>(rx, ry, rz, [&](int lm, double y, double yx, double yy, double yz) {
                const int l = lm_l2(lm);
                double h0 = 0.0, h1 = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double mv = Mrow[lm * n + k];
                    h0 += r0[k * L1 + l] * mv;
                    h1 += r1[k * L1 + l] * mv;
                }
                const double yh1 = y * h1;
                gx += rx * yh1 + yx * h0;
                gy += ry * yh1 + yy * h0;
                gz += rz * yh1 + yz * h0;
            });
            const int p = ordered ? ord.rev[e] : -1;
            if (ordered) {                              // gathered through the reverse edge
                slots[3 * e] = gx;
                slots[3 * e + 1] = gy;
                slots[3 * e + 2] = gz;
            }
            if (p < 0) {
                atomicAdd(&grad[3 * (size_t)a], gx);
                atomicAdd(&grad[3 * (size_t)a + 1], gy);
                atomicAdd(&grad[3 * (size_t)a + 2], gz);
            }
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
            if (ordered && ca == c) {
                ord.att[3 * (size_t)c] = sx;
                ord.att[3 * (size_t)c + 1] = sy;
                ord.att[3 * (size_t)c + 2] = sz;
            } else {
                atomicAdd(&grad[3 * (size_t)ca], -sx);
                atomicAdd(&grad[3 * (size_t)ca + 1], -sy);
                atomicAdd(&grad[3 * (size_t)ca + 2], -sz);
            }
        }
    }
}

template <int LMAX, typename TM, int RAD, int NN>
int launch_vjp_gto(const cusoap_basis_t& b, const double* pos, int n_atoms, const int32_t* atom_sp,
                   const double* shifts, const double* centers, int n_centers, const int32_t* center_attach,
                   const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* M,
                   const int32_t* order, double* grad, const VjpOrder& ord, hipStream_t st) {
    const size_t cols = sizeof(TM) * (NN > 0 ? 0 : 2 * 256 * (size_t)b.n_max) +
                        (RAD == kRadMixed && NN == 0 ? sizeof(double) * 256 * (size_t)b.n_max : 0);
    const size_t per_bytes = sizeof(TM) * (size_t)b.n_species * b.n_max * (LMAX + 1) * (LMAX + 1);
    // as many centers per workgroup as fit in 64 KB of coefficient rows (at least one); the
    // register-table form keeps them within 32 KB, so that three workgroups with their hash
    // tables share a CU (H/C/N/O, 6.3 KB rows: 10 centers took 0.48 ms against 0.41 for 4)
    const size_t rows_budget = NN > 0 ? 32 * 1024 : 64 * 1024;
    int ngrp = (int)std::min<size_t>(vjp_group<NN>(), std::max<size_t>(1, (rows_budget - std::min<size_t>(cols, rows_budget)) / per_bytes));
    const size_t smem = cols + ngrp * per_bytes;
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: coefficient block too large for LDS");
    auto kern = vjp_gto_kernel<LMAX, TM, RAD, NN>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("vjp_edges", st);
    const int groups = (n_centers + ngrp - 1) / ngrp, per_xcd = (groups + 7) / 8;
    kern<<<8 * per_xcd, 256, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, center_attach, center_species,
                                         offsets, nbr, (const TM*)M, order, n_centers, ngrp, per_xcd, grad, ord);
    prof_end("vjp_edges", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <int LMAX>
int launch_vjp_edges(const cusoap_basis_t& b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                     const double* shifts, const double* centers, int n_centers, const int32_t* center_attach,
                     const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* M,
                     const int32_t* order, double* grad, const VjpOrder& ord, hipStream_t st) {
#define CUSOAP_VJP_ARGS b, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach, center_species, offsets, nbr, M, order, grad, ord, st
    if (b.rad_table) {
        if (prec == CUSOAP_PREC_F64) return launch_vjp_gto<LMAX, double, kRadTable, 0>(CUSOAP_VJP_ARGS);
        // registers for the radial vectors (float table form) at n_max 4..8
        switch (b.n_max) {
            case 8: return launch_vjp_gto<LMAX, float, kRadTable, 8>(CUSOAP_VJP_ARGS);
            case 7: return launch_vjp_gto<LMAX, float, kRadTable, 7>(CUSOAP_VJP_ARGS);
            case 6: return launch_vjp_gto<LMAX, float, kRadTable, 6>(CUSOAP_VJP_ARGS);
            case 5: return launch_vjp_gto<LMAX, float, kRadTable, 5>(CUSOAP_VJP_ARGS);
            case 4: return launch_vjp_gto<LMAX, float, kRadTable, 4>(CUSOAP_VJP_ARGS);
            default: return launch_vjp_gto<LMAX, float, kRadTable, 0>(CUSOAP_VJP_ARGS);
        }
    }
    if (b.rbf == CUSOAP_RBF_GTO) {
        if (prec == CUSOAP_PREC_F64) return launch_vjp_gto<LMAX, double, kRadPrim, 0>(CUSOAP_VJP_ARGS);
        // (the register variant would exceed 256 VGPRs on these paths)
        if (prec == CUSOAP_PREC_F32) return launch_vjp_gto<LMAX, float, kRadPrim, 0>(CUSOAP_VJP_ARGS);
        return launch_vjp_gto<LMAX, float, kRadMixed, 0>(CUSOAP_VJP_ARGS);
    }
#undef CUSOAP_VJP_ARGS
    const int L2 = (LMAX + 1) * (LMAX + 1);
    const size_t smem = sizeof(double) * ((size_t)b.n_species * b.n_max * L2 + 2 * 64 * (size_t)b.n_max * (LMAX + 1));
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: LDS tile too large");
    auto kern = vjp_poly_kernel<LMAX>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("vjp_edges", st);
    kern<<<n_centers, 64, smem, st>>>(b, pos, n_atoms, atom_sp, shifts, centers, center_attach, offsets, nbr,
                                      (const double*)M, order, n_centers, grad, ord);
    prof_end("vjp_edges", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <typename TG>
int launch_ps_mfma(const cusoap_basis_t& b, const void* coeffs, const int32_t* pair_feature, const void* gd,
                   int fold_k, int n_centers, void* M, void* Wbuf, size_t smem, hipStream_t st) {
    const int L1 = b.l_max + 1, SN = b.n_species * b.n_max;
    const int nw = psw_size(L1, SN);
    prof_begin("vjp_adjoint", st);
    ps_weights_kernel<TG, float><<<(nw + 255) / 256, 256, 0, st>>>(L1, SN, pair_feature, (const TG*)gd, (float*)Wbuf);
    CUSOAP_LAUNCH_CHECK();
    auto kern = ps_adjoint_mfma_kernel;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    // one resident round of workgroups (every wavefront takes the same number of centers)
    int dev = 0, n_cu = 0, per_cu = 0;
    CUSOAP_HIP(hipGetDevice(&dev));
    CUSOAP_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    CUSOAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, 256, smem));
    const int grid = std::max(1, std::min((n_centers + 3) / 4, std::max(1, n_cu * per_cu)));
    kern<<<grid, 256, smem, st>>>(b.n_species, b.n_max, L1, (const float*)coeffs, (const float*)Wbuf, fold_k,
                                  n_centers, (float*)M);
    prof_end("vjp_adjoint", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

template <typename TC, typename TG, typename TW, typename TM>
int launch_ps_shared(const cusoap_basis_t& b, const void* coeffs, const int32_t* pair_feature, const void* gd,
                     int fold_beta, int fold_k, int n_centers, void* M, void* Wbuf, size_t smem, hipStream_t st) {
    const int L1 = b.l_max + 1, SN = b.n_species * b.n_max;
    const int nw = psw_size(L1, SN);
    prof_begin("vjp_adjoint", st);
    ps_weights_kernel<TG, TW><<<(nw + 255) / 256, 256, 0, st>>>(L1, SN, pair_feature, (const TG*)gd, (TW*)Wbuf);
    CUSOAP_LAUNCH_CHECK();
    auto kern = ps_adjoint_shared_kernel<TC, TW, TM>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    const int grid = n_centers < 2048 ? n_centers : 2048;
    kern<<<grid, 256, smem, st>>>(b.n_species, b.n_max, L1, (const TC*)coeffs, (const TW*)Wbuf, b.gto_beta, fold_beta,
                                  fold_k, n_centers, (TM*)M);
    prof_end("vjp_adjoint", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

// bytes of M (sized for float64) at the head of the vjp workspace; W follows
size_t vjp_m_bytes(const cusoap_basis_t* b, int n_centers) {
    const int L2 = (b->l_max + 1) * (b->l_max + 1);
    return align_up(sizeof(double) * (size_t)std::max(n_centers, 1) * b->n_species * b->n_max * L2, 256);
}

size_t vjp_w_bytes(const cusoap_basis_t* b) {
    return align_up(sizeof(double) * (size_t)psw_size(b->l_max + 1, b->n_species * b->n_max), 256);
}

template <typename TC, typename TG, typename TM>
int launch_adjoint(const cusoap_basis_t& b, int mode, const void* coeffs, const int32_t* pair_feature, const void* gd,
                   int gd_broadcast, int n_feat, int fold_beta, int fold_k, int n_centers, void* M, hipStream_t st) {
    const int L1 = b.l_max + 1, SN = b.n_species * b.n_max;
    const int per = SN * L1 * L1;
    if (mode == CUSOAP_MODE_PS && gd_broadcast) {
        // one weight matrix for all centers: the grid-stride kernel when it fits in LDS
        const bool w64 = fold_beta || sizeof(TM) == 8;
        const size_t tw = w64 ? 8 : 4;
        const size_t smem = tw * ((size_t)psw_size(L1, SN) + 2 * (size_t)per + (fold_beta ? (size_t)L1 * b.n_max * b.n_max : 0));
        void* Wbuf = (char*)M + vjp_m_bytes(&b, n_centers);
        if constexpr (std::is_same<TC, float>::value && std::is_same<TM, float>::value) {
            if (!fold_beta && per <= 64 * kPsMfmaPre && SN <= 16 * kPsMfmaMaxNT && ps_mfma_enabled()) {
                const size_t smem_m = sizeof(float) * ((size_t)((psw_size(L1, SN) + 3) & ~3) + ((L1 * L1 + 3) & ~3) +
                                                       8 * (size_t)ps_mfma_slice(per));
                if (smem_m <= 160 * 1024)
                    return launch_ps_mfma<TG>(b, coeffs, pair_feature, gd, fold_k, n_centers, M, Wbuf, smem_m, st);
            }
        }
        if (smem <= 160 * 1024 && per <= 256 * kPsPre)
            return w64 ? launch_ps_shared<TC, TG, double, TM>(b, coeffs, pair_feature, gd, fold_beta, fold_k, n_centers, M, Wbuf, smem, st)
                       : launch_ps_shared<TC, TG, float, TM>(b, coeffs, pair_feature, gd, fold_beta, fold_k, n_centers, M, Wbuf, smem, st);
    }
    const size_t smem = sizeof(double) * (2 * (size_t)per + n_feat + (mode == CUSOAP_MODE_PS ? (size_t)L1 * SN * (SN + 1) : 0) +
                                          (fold_beta ? (size_t)L1 * b.n_max * b.n_max : 0));
    if (smem > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: descriptor row too large for LDS");
    auto kern = coeff_adjoint_kernel<TC, TG, TM>;
    if (smem > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    prof_begin("vjp_adjoint", st);
    kern<<<n_centers, 256, smem, st>>>(b, mode, (const TC*)coeffs, pair_feature, (const TG*)gd, gd_broadcast, n_feat,
                                       fold_beta, fold_k, (TM*)M);
    prof_end("vjp_adjoint", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace

// build.py compiles this file once per l_max value with -DCUSOAP_ONLY_L=<l> (the edge
// kernels of one degree) and once without it (everything else, including the entry
// points, which call the per-degree functions).
#define CUSOAP_VJP_PARAMS const cusoap_basis_t& b, int prec, const double* pos, int n_atoms, const int32_t* atom_sp, \
    const double* shifts, const double* centers, int n_centers, const int32_t* center_attach,                     \
    const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* M,                       \
    const int32_t* order, double* grad, const VjpOrder& ord, hipStream_t st
#define CUSOAP_VJP_DECL(L) int vjp_edges_l##L(CUSOAP_VJP_PARAMS);
CUSOAP_VJP_DECL(0) CUSOAP_VJP_DECL(1) CUSOAP_VJP_DECL(2) CUSOAP_VJP_DECL(3) CUSOAP_VJP_DECL(4)
CUSOAP_VJP_DECL(5) CUSOAP_VJP_DECL(6) CUSOAP_VJP_DECL(7) CUSOAP_VJP_DECL(8) CUSOAP_VJP_DECL(9)
#undef CUSOAP_VJP_DECL

#ifdef CUSOAP_ONLY_L
#define CUSOAP_VJP_DEF(L)                                                                                        \
    int vjp_edges_l##L(CUSOAP_VJP_PARAMS) {                                                                      \
        return launch_vjp_edges<L>(b, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach,     \
                                   center_species, offsets, nbr, M, order, grad, ord, st);                       \
    }
#define CUSOAP_VJP_DEF2(L) CUSOAP_VJP_DEF(L)
CUSOAP_VJP_DEF2(CUSOAP_ONLY_L)
#else
size_t center_order_ws_bytes(int n);                        // order.hip
int center_order(const double* centers, int n, void* ws, int32_t* order, hipStream_t st);

// dL/dc rows (M), the shared weight matrix (W), then the centers' Morton order and the
// workspace that sorts it
size_t vjp_core_ws_bytes(const cusoap_basis_t* b, int n_centers) { return vjp_m_bytes(b, n_centers) + vjp_w_bytes(b); }
size_t vjp_order_bytes(int n_centers) {
    return align_up(sizeof(int32_t) * (size_t)std::max(n_centers, 1), 256) + center_order_ws_bytes(n_centers);
}
size_t vjp_ws_bytes(const cusoap_basis_t* b, int n_centers) {
    return vjp_core_ws_bytes(b, n_centers) + vjp_order_bytes(n_centers);
}

namespace {

// grad = 0 by a kernel rather than hipMemsetAsync: the allocation-free entry point is
// meant for callers that capture HIP graphs, where memset nodes are avoided
__global__ void zero_f64_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

int zero_f64(double* p, int64_t n, hipStream_t st) {
    if (n <= 0) return CUSOAP_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    zero_f64_kernel<<<(int)blocks, 256, 0, st>>>(p, n);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

__device__ __forceinline__ bool same_point(const double* __restrict__ u, const double* __restrict__ v) {
    return u[0] == v[0] && u[1] == v[1] && u[2] == v[2];
}

__global__ void fill_rev_kernel(const int64_t* __restrict__ offsets, int n_centers, int32_t* __restrict__ rev) {
    const int64_t n = offsets[n_centers];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        rev[i] = -1;
}

// Per-center search block of the reverse-edge pass, 16 int32 (64 bytes): range start,
// length, 1 if the center is its own atom (centers[c] == positions[c] bitwise), then the
// keys at positions floor(k len / 8), k = 0..7, of its sorted range (INT_MAX past the end)
constexpr int kRevBlock = 16;

__global__ void reverse_index_kernel(const double* __restrict__ pos, int n_atoms, const double* __restrict__ centers,
                                     int n_centers, const int64_t* __restrict__ offsets,
                                     const int32_t* __restrict__ nbr, int32_t* __restrict__ index) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_centers) return;
    const int64_t s0 = offsets[c];
    const int len = (int)(offsets[c + 1] - s0);
    int32_t* blk = index + (size_t)kRevBlock * c;
    blk[0] = (int32_t)s0;
    blk[1] = len;
    blk[2] = (c < n_atoms && same_point(centers + 3 * (size_t)c, pos + 3 * (size_t)c)) ? 1 : 0;
    blk[3] = 0;
    for (int k = 0; k < 8; ++k) blk[4 + k] = len > 0 ? nbr[s0 + (int64_t)k * len / 8] : INT_MAX;
    for (int k = 12; k < kRevBlock; ++k) blk[k] = 0;
}

// rev for the edges of center c (one wavefront per center; consecutive workgroups go to
// different XCDs, each XCD taking a contiguous range of centers, whose neighbours' ranges
// then share its L2): the edge c -> a and its reverse a -> c (the image with the opposite
// shift) are linked both ways by the one of the two with c < a (c == a: the image with the
// lower index), which finds the other in a's range: a's search block (one read) gives the
// eighth of the range that can hold the key, then a binary search within it.  rev stays -1
// (set beforehand) when the reverse is missing or the lists are not the atoms' own
// (centers[c] != positions[c] bitwise).  The image with the opposite shift mirrors img in
// a -n..n grid of shifts; that guess is checked exactly, then every image is tried.
__global__ __launch_bounds__(256) void edge_reverse_kernel(int n_atoms, const double* __restrict__ shifts, int n_img,
                                                           int n_centers, int per_xcd,
                                                           const int64_t* __restrict__ offsets,
                                                           const int32_t* __restrict__ nbr,
                                                           const int32_t* __restrict__ index,
                                                           int32_t* __restrict__ rev) {
    const int lane = threadIdx.x & 63;
    const int c = ((int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3)) * 4 + (threadIdx.x >> 6);
    if (c >= n_centers || c >= n_atoms) return;
    if (!index[(size_t)kRevBlock * c + 2]) return;
    const double inv_na = 1.0 / (double)n_atoms;
    for (int64_t e = offsets[c] + lane; e < offsets[c + 1]; e += 64) {
        int img, a;
        split_ext(nbr[e], n_atoms, inv_na, img, a);
        if (a < c || a >= n_centers) continue;
        const double* sh = shifts + 3 * (size_t)img;
        auto opposite = [&](int t) {
            return shifts[3 * t] == -sh[0] && shifts[3 * t + 1] == -sh[1] && shifts[3 * t + 2] == -sh[2];
        };
        int t = n_img - 1 - img;
        if (!opposite(t)) {
            t = -1;
            for (int k = 0; k < n_img && t < 0; ++k)
                if (opposite(k)) t = k;
        }
        if (t < 0 || (a == c && t > img)) continue;
        if (a == c && t == img) {                  // its own reverse (the zero shift)
            rev[e] = (int32_t)e;
            continue;
        }
        const int4* blk = (const int4*)(index + (size_t)kRevBlock * a);
        const int4 h = blk[0], s0 = blk[1], s1 = blk[2];
        if (!h.z) continue;
        const int key = t * n_atoms + c;
        // the last sample <= key: its eighth of the range holds the key if the range does
        const int j = (s0.x <= key) + (s0.y <= key) + (s0.z <= key) + (s0.w <= key) + (s1.x <= key) +
                      (s1.y <= key) + (s1.z <= key) + (s1.w <= key) - 1;
        if (j < 0) continue;
        int lo = h.x + (int)((int64_t)j * h.y / 8);
        const int end = h.x + (int)((int64_t)(j + 1) * h.y / 8);
        int hi = end;
        while (lo < hi) {
            const int mid = lo + ((hi - lo) >> 1);
            if (nbr[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        if (lo < end && nbr[lo] == key) {
            rev[e] = (int32_t)lo;
            rev[lo] = (int32_t)e;
        }
    }
}

// grad[a] += the contributions of the edges c -> a (rev of the entries a -> c of a's own
// range), summed in float64 in a fixed order (lane-strided, then a butterfly over the
// wavefront), minus the sum of its own center's edges when that center moves with it.
// One wavefront per atom.
template <typename TM>
__global__ __launch_bounds__(256) void vjp_gather_kernel(const int64_t* __restrict__ offsets, int n_centers,
                                                         const int32_t* __restrict__ center_attach, VjpOrder ord,
                                                         double* __restrict__ grad) {
    if (!vjp_ordered(ord, offsets, n_centers)) return;
    const int lane = threadIdx.x & 63;
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= n_centers) return;
    const TM* sl = (const TM*)ord.slots;
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t p = offsets[a] + lane; p < offsets[a + 1]; p += 64) {
        const int q = ord.rev[p];
        if (q >= 0)
#pragma unroll
            for (int d = 0; d < 3; ++d) v[d] += (double)sl[3 * (int64_t)q + d];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] += __shfl_xor(v[d], o, 64);
    if (lane == 0) {
        const bool moves = center_attach && center_attach[a] == a;
#pragma unroll
        for (int d = 0; d < 3; ++d) grad[3 * (size_t)a + d] += moves ? v[d] - ord.att[3 * (size_t)a + d] : v[d];
    }
}

// bytes of the fixed-order sums beyond the dL/dc rows: per center (attach sums) and per edge
// (the three contributions in the edge pass's type)
size_t order_center_bytes(int n_centers) { return align_up(sizeof(double) * 3 * (size_t)std::max(n_centers, 1), 256); }
size_t order_edge_bytes(int prec) { return 3 * (prec == CUSOAP_PREC_F64 ? sizeof(double) : sizeof(float)); }

// checks shared by both entry points; returns CUSOAP_OK or the error, and the precision
// the kernels run in (the polynomial basis always runs in float64)
int vjp_check(const cusoap_basis_t* b, int mode, int& prec, int n_centers, const int32_t* center_species,
              const int32_t* pair_feature) {
    if (!b || b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: n_max must be 1..16 and l_max 0..9");
    if (mode == CUSOAP_MODE_PS && !pair_feature) return fail(CUSOAP_ERR_ARG, "vjp: pair_feature table missing");
    if (prec < 0 || prec > 2) return fail(CUSOAP_ERR_ARG, "vjp: unknown precision");
    if (n_centers < 0) return fail(CUSOAP_ERR_ARG, "vjp: negative center count");
    if (n_centers > 0 && b->rad_table && (!center_species || b->rad_intervals < 1 || !(b->rad_rmax > 0.0)))
        return fail(CUSOAP_ERR_ARG, "vjp: radial table needs center_species and a valid grid");
    if (b->rbf != CUSOAP_RBF_GTO && !b->rad_table) prec = CUSOAP_PREC_F64;
    return CUSOAP_OK;
}

// dL/dc (M, then W for a shared gradient, in ws) and the edge pass of centers
// [0, n_centers) of the given arrays, adding into grad (not zeroed here)
int vjp_core(const cusoap_basis_t* b, int mode, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
             const double* shifts, const double* centers, int n_centers, const int32_t* center_attach,
             const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* coeffs,
             int coeffs_dtype, const int32_t* pair_feature, const void* gd, int gd_dtype, int gd_broadcast,
             int n_feat, void* ws, const int32_t* order, hipEvent_t order_ready, double* grad, const VjpOrder& ord,
             hipStream_t st) {
    const int fold = (b->rbf == CUSOAP_RBF_GTO && prec != CUSOAP_PREC_MIXED && !b->rad_table) ? 1 : 0;
    const int fold_k = (b->rbf == CUSOAP_RBF_GTO || b->rad_table) ? 1 : 0;   // the GTO / table edge kernels
    int rc;
#define CUSOAP_ADJ(TC, TG, TM) \
    launch_adjoint<TC, TG, TM>(*b, mode, coeffs, pair_feature, gd, gd_broadcast, n_feat, fold, fold_k, n_centers, ws, st)
    const bool c64 = coeffs_dtype == CUSOAP_DT_F64, g64 = gd_dtype == CUSOAP_DT_F64;
    if (prec == CUSOAP_PREC_F64) {
        if (c64) rc = g64 ? CUSOAP_ADJ(double, double, double) : CUSOAP_ADJ(double, float, double);
        else rc = g64 ? CUSOAP_ADJ(float, double, double) : CUSOAP_ADJ(float, float, double);
    } else {
        if (c64) rc = g64 ? CUSOAP_ADJ(double, double, float) : CUSOAP_ADJ(double, float, float);
        else rc = g64 ? CUSOAP_ADJ(float, double, float) : CUSOAP_ADJ(float, float, float);
    }
#undef CUSOAP_ADJ
    if (rc) return rc;
    if (order_ready) CUSOAP_HIP(hipStreamWaitEvent(st, order_ready, 0));
#define CUSOAP_CASE(L) \
    case L: return vjp_edges_l##L(*b, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach, center_species, offsets, nbr, ws, order, grad, ord, st);
    switch (b->l_max) {
        CUSOAP_CASE(0) CUSOAP_CASE(1) CUSOAP_CASE(2) CUSOAP_CASE(3) CUSOAP_CASE(4)
        CUSOAP_CASE(5) CUSOAP_CASE(6) CUSOAP_CASE(7) CUSOAP_CASE(8) CUSOAP_CASE(9)
        default: break;
    }
#undef CUSOAP_CASE
    return fail(CUSOAP_ERR_UNSUPPORTED, "vjp: l_max out of range");
}

// Workspace of the allocation-free entry point: a file-scope device array, one per GPU
// (each device holds its own copy at its own address until the process ends).  In host
// code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed to kernels.
constexpr size_t kNoAllocScratchBytes = CUSOAP_VJP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;      // device ordinal -> device address of `scratch`

int scratch_address(void** out) {
    int dev = 0;
    CUSOAP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CUSOAP_ERR_HIP, std::string("vjp_noalloc: hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
        it = g_scratch_ptr.emplace(dev, p).first;
    }
    *out = it->second;
    return CUSOAP_OK;
}

}  // namespace

// A second stream per device (default priority, non-blocking) for the centers' Morton
// order: it depends on the coordinates only, so its small kernels run beside the
// adjoint's first kernel; the edge pass waits on an event.
std::mutex g_side_mu;
std::map<int, hipStream_t> g_side;

int side_stream(hipStream_t* out) {
    int dev = 0;
    CUSOAP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_side_mu);
    auto it = g_side.find(dev);
    if (it == g_side.end()) {
        hipStream_t s = nullptr;
        CUSOAP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        it = g_side.emplace(dev, s).first;
    }
    *out = it->second;
    return CUSOAP_OK;
}

// the centers' Morton order, in the workspace after the dL/dc rows (see vjp_ws_bytes), on
// the side stream after the work already on st; *ready is recorded after whatever was
// enqueued there, also when a launch failed (the caller joins st to it on every path
// before returning, see join_side, so no side work outlives the call's workspace)
int morton_order(const cusoap_basis_t* b, const double* centers, int n_centers, void* ws, int32_t** order,
                 hipEvent_t* ready, hipStream_t st) {
    char* p = (char*)ws + vjp_core_ws_bytes(b, n_centers);
    *order = (int32_t*)p;
    *ready = nullptr;
    hipStream_t side = nullptr;
    int rc = side_stream(&side);
    if (rc) return rc;
    hipEvent_t fork = nullptr;
    CUSOAP_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    CUSOAP_HIP(hipEventRecord(fork, st));
    const hipError_t we = hipStreamWaitEvent(side, fork, 0);
    (void)hipEventDestroy(fork);             // released once recorded work completes
    if (we != hipSuccess) return fail(CUSOAP_ERR_HIP, std::string("vjp: stream wait: ") + hipGetErrorString(we));
    rc = center_order(centers, n_centers, p + align_up(sizeof(int32_t) * (size_t)n_centers, 256), *order, side);
    CUSOAP_HIP(hipEventCreateWithFlags(ready, hipEventDisableTiming));
    CUSOAP_HIP(hipEventRecord(*ready, side));
    return rc;
}

// st waits for the side stream's order work (a second wait after the edge pass's is
// harmless), then the event is released
void join_side(hipEvent_t ready, hipStream_t st) {
    if (!ready) return;
    (void)hipStreamWaitEvent(st, ready, 0);
    (void)hipEventDestroy(ready);
}

int vjp_impl(const cusoap_basis_t* b, int mode, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
             const double* shifts, int n_img, const double* centers, int n_centers, const int32_t* center_attach,
             const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const void* coeffs, int coeffs_dtype,
             const int32_t* pair_feature, const void* gd, int gd_dtype, int gd_broadcast, int n_feat, void* ws,
             size_t ws_bytes, double* grad, hipStream_t st) {
    (void)n_img;
    int rc = vjp_check(b, mode, prec, n_centers, center_species, pair_feature);
    if (rc) return rc;
    if (ws_bytes < vjp_ws_bytes(b, n_centers)) return fail(CUSOAP_ERR_WORKSPACE, "vjp: workspace too small");
    if ((rc = zero_f64(grad, 3 * (int64_t)n_atoms, st))) return rc;
    if (n_centers == 0) return CUSOAP_OK;
    int32_t* order = nullptr;
    hipEvent_t ready = nullptr;
    if ((rc = morton_order(b, centers, n_centers, ws, &order, &ready, st))) {
        join_side(ready, st);
        return rc;
    }
    rc = vjp_core(b, mode, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach, center_species,
                  offsets, nbr, coeffs, coeffs_dtype, pair_feature, gd, gd_dtype, gd_broadcast, n_feat, ws, order, ready,
                  grad, VjpOrder{nullptr, nullptr, nullptr, 0}, st);
    join_side(ready, st);
    return rc;
}

size_t vjp_ordered_ws_bytes(const cusoap_basis_t* b, int prec, int n_centers, int64_t n_edges) {
    if (b->rbf != CUSOAP_RBF_GTO && !b->rad_table) prec = CUSOAP_PREC_F64;
    const size_t e = (size_t)std::max<int64_t>(n_edges, 1);
    return vjp_ws_bytes(b, n_centers) + order_center_bytes(n_centers) + align_up(order_edge_bytes(prec) * e, 256);
}

size_t edge_reverse_ws_bytes(int n_centers) {
    return align_up(sizeof(int32_t) * kRevBlock * (size_t)std::max(n_centers, 1), 256);
}

int edge_reverse_impl(const double* pos, int n_atoms, const double* shifts, int n_img, const double* centers,
                      int n_centers, const int64_t* offsets, const int32_t* nbr, void* ws, size_t ws_bytes,
                      int32_t* rev, hipStream_t st) {
    if (n_atoms < 1 || n_img < 1 || n_centers < 0) return fail(CUSOAP_ERR_ARG, "edge_reverse: bad sizes");
    if (!pos || !shifts || !centers || !offsets || !nbr || !rev || !ws)
        return fail(CUSOAP_ERR_ARG, "edge_reverse: null argument");
    if (ws_bytes < edge_reverse_ws_bytes(n_centers)) return fail(CUSOAP_ERR_WORKSPACE, "edge_reverse: workspace too small");
    if (n_centers == 0) return CUSOAP_OK;
    int32_t* index = (int32_t*)ws;
    // the edge count is read on the device (offsets[n_centers]): -1 over the whole range
    // by a grid-stride fill bounded by the last offset
    prof_begin("edge_reverse", st);
    fill_rev_kernel<<<1024, 256, 0, st>>>(offsets, n_centers, rev);
    CUSOAP_LAUNCH_CHECK();
    reverse_index_kernel<<<(n_centers + 255) / 256, 256, 0, st>>>(pos, n_atoms, centers, n_centers, offsets, nbr, index);
    CUSOAP_LAUNCH_CHECK();
    const int groups = (n_centers + 3) / 4, per = (groups + 7) / 8;
    edge_reverse_kernel<<<8 * per, 256, 0, st>>>(n_atoms, shifts, n_img, n_centers, per, offsets, nbr, index, rev);
    prof_end("edge_reverse", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

int vjp_ordered_impl(const cusoap_basis_t* b, int mode, int prec, const double* pos, int n_atoms, const int32_t* atom_sp,
                     const double* shifts, int n_img, const double* centers, int n_centers, const int32_t* center_attach,
                     const int32_t* center_species, const int64_t* offsets, const int32_t* nbr, const int32_t* rev,
                     int64_t n_edges, const void* coeffs, int coeffs_dtype, const int32_t* pair_feature, const void* gd,
                     int gd_dtype, int gd_broadcast, int n_feat, void* ws, size_t ws_bytes, double* grad, hipStream_t st) {
    (void)n_img;
    int rc = vjp_check(b, mode, prec, n_centers, center_species, pair_feature);
    if (rc) return rc;
    if (n_centers != n_atoms) return fail(CUSOAP_ERR_ARG, "vjp_ordered: every atom must be its own center");
    if (!rev) return fail(CUSOAP_ERR_ARG, "vjp_ordered: reverse-edge index missing");
    if (n_edges < 0 || n_edges > INT32_MAX) return fail(CUSOAP_ERR_ARG, "vjp_ordered: edge count out of range");
    const size_t base = vjp_ws_bytes(b, n_centers);
    if (ws_bytes < vjp_ordered_ws_bytes(b, prec, n_centers, n_edges))
        return fail(CUSOAP_ERR_WORKSPACE, "vjp_ordered: workspace too small");
    if ((rc = zero_f64(grad, 3 * (int64_t)n_atoms, st))) return rc;
    if (n_centers == 0) return CUSOAP_OK;
    int32_t* order = nullptr;
    hipEvent_t ready = nullptr;
    if ((rc = morton_order(b, centers, n_centers, ws, &order, &ready, st))) {
        join_side(ready, st);
        return rc;
    }
    char* p = (char*)ws + base;
    const VjpOrder ord{rev, p + order_center_bytes(n_centers), (double*)p, n_edges};
    rc = vjp_core(b, mode, prec, pos, n_atoms, atom_sp, shifts, centers, n_centers, center_attach, center_species,
                  offsets, nbr, coeffs, coeffs_dtype, pair_feature, gd, gd_dtype, gd_broadcast, n_feat, ws, order, ready,
                  grad, ord, st);
    join_side(ready, st);
    if (rc) return rc;
    prof_begin("vjp_gather", st);
    if (prec == CUSOAP_PREC_F64)
        vjp_gather_kernel<double><<<(n_centers + 3) / 4, 256, 0, st>>>(offsets, n_centers, center_attach, ord, grad);
    else
        vjp_gather_kernel<float><<<(n_centers + 3) / 4, 256, 0, st>>>(offsets, n_centers, center_attach, ord, grad);
    prof_end("vjp_gather", st);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

// Centers in chunks whose workspace fits the static array; the chunks add into grad.
// Nothing in the array needs clearing: M and W are written before they are read.
int vjp_noalloc_impl(const cusoap_basis_t* b, int mode, int prec, const double* pos, int n_atoms,
                     const int32_t* atom_sp, const double* shifts, int n_img, const double* centers, int n_centers,
                     const int32_t* center_attach, const int32_t* center_species, const int64_t* offsets,
                     const int32_t* nbr, const void* coeffs, int coeffs_dtype, const int32_t* pair_feature,
                     const void* gd, int gd_dtype, int gd_broadcast, int n_feat, double* grad, hipStream_t st) {
    (void)n_img;
    int rc = vjp_check(b, mode, prec, n_centers, center_species, pair_feature);
    if (rc) return rc;
    const size_t per_center = vjp_m_bytes(b, 1);     // one center's rows, rounded up to 256 bytes
    const size_t wbytes = vjp_w_bytes(b);
    if (per_center + wbytes > kNoAllocScratchBytes)
        return fail(CUSOAP_ERR_CAPACITY, "vjp_noalloc: one center's adjoint rows exceed the 64 MiB scratch");
    const int L2 = (b->l_max + 1) * (b->l_max + 1);
    const size_t row_f64 = sizeof(double) * (size_t)b->n_species * b->n_max * L2;
    int chunk = (int)std::min<size_t>((kNoAllocScratchBytes - wbytes - 256) / row_f64, (size_t)INT32_MAX);
    while (chunk > 1 && vjp_core_ws_bytes(b, chunk) > kNoAllocScratchBytes) --chunk;
    if (chunk < 1 || vjp_core_ws_bytes(b, chunk) > kNoAllocScratchBytes)
        return fail(CUSOAP_ERR_CAPACITY, "vjp_noalloc: basis too large for the 64 MiB scratch");
    void* ws = nullptr;
    if ((rc = scratch_address(&ws))) return rc;
    if ((rc = zero_f64(grad, 3 * (int64_t)n_atoms, st))) return rc;
    const size_t csz = coeffs_dtype == CUSOAP_DT_F64 ? 8 : 4, gsz = gd_dtype == CUSOAP_DT_F64 ? 8 : 4;
    const size_t per = (size_t)b->n_species * b->n_max * L2;
    for (int c0 = 0; c0 < n_centers; c0 += chunk) {
        const int nc = std::min(chunk, n_centers - c0);
        rc = vjp_core(b, mode, prec, pos, n_atoms, atom_sp, shifts, centers + 3 * (size_t)c0, nc,
                      center_attach ? center_attach + c0 : nullptr, center_species ? center_species + c0 : nullptr,
                      offsets + c0, nbr, (const char*)coeffs + csz * per * c0, coeffs_dtype, pair_feature,
                      gd_broadcast ? gd : (const void*)((const char*)gd + gsz * (size_t)n_feat * c0), gd_dtype,
                      gd_broadcast, n_feat, ws, nullptr, nullptr, grad, VjpOrder{nullptr, nullptr, nullptr, 0}, st);
        if (rc) return rc;
    }
    return CUSOAP_OK;
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
#endif  // CUSOAP_ONLY_L

}  // namespace cusoap
