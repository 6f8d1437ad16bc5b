// Learned radial model: one small network per species pair approximating the
// orthonormal GTO radial functions B_nl(r) and their slopes, trained on the
// device, and the cubic-Hermite table of the radial model that the coefficient
// and adjoint kernels interpolate.  No reference counterpart: the reference
// evaluates the exact radial integral per edge (soap_generator.py:1358-1391,
// 2389-2405).
//
// Network (per pair, float32): x in [-1, 1] -> 64 tanh -> 64 tanh -> O linear,
// O = n_max (l_max+1), outputs l-major (j = l n_max + n).  Parameter vector
// per pair: W1[64] b1[64] W2[64][64] b2[64] W3[O][64] b3[O].  The outputs model
// the radial projections g_nl(r) = (r / r_fit)^l B_nl(r), scaled per output: the
// function the reference multiplies with the unit-vector harmonics, up to the
// constant r_fit^l.  Its error, not that of B_nl, is what reaches the coefficients.
// Training kernels:
//   radial_exact_kernel   exact float64 B_nl(r), dB_nl/dr (training targets, table fallback)
//   mlp_eval_kernel       y and dy/dx of every pair at a batch of inputs
//   mlp_grad_kernel       loss and per-tile gradient of
//                         1/(O B) sum [wv_o (y - t)^2 + wd_o (dy/dx - dt)^2] over a
//                         hashed minibatch; value and tangent paths reversed by hand
//   mlp_reduce_kernel     fixed-order sum of the tile gradients (deterministic)
//   adam_kernel           Adam with bias correction
//   mlp_rows_kernel       hidden features [h2, 1, t] / [dh2/dx, 0, dt] of every sample
//   mlp_normal_kernel     float64 normal equations of the output layer (chunk partials)
//   mlp_solve_kernel      Cholesky solve per output channel -> W3, b3 (exact refit)
//   table_kernel          B(r_i), dB/dr(r_i) at the table nodes: B = g / u^l and
//                         dB/dr = g'/u^l - l g / (r_fit u^(l+1)), u = r_i / r_fit, from the
//                         network where the pair is enabled and 0 < r_i <= r_fit, else the
//                         exact function (r_i = 0 always: the exact self term's radius).  At small r_i the quotient
//                         carries the network's error divided by u^l; the kernels multiply
//                         B by r^l Y_lm, so the error of B r^l stays that of g.
#include <cmath>

#include "common.hpp"

namespace cusoap {

namespace {

constexpr int kH = 64;
constexpr int kHP = kH + 1;
constexpr int kT = 32;          // samples per tile
constexpr int kThr = 256;
constexpr int kMaxO = 128;

struct MlpOff {
    int W1, b1, W2, b2, W3, b3, P;
};
__host__ __device__ inline MlpOff mlp_off(int O) {
    MlpOff o;
    o.W1 = 0;
    o.b1 = kH;
    o.W2 = 2 * kH;
    o.b2 = 2 * kH + kH * kH;
    o.W3 = 3 * kH + kH * kH;
    o.b3 = o.W3 + O * kH;
    o.P = o.b3 + O;
    return o;
}

// LDS image of one pair's network plus the activations of one tile
struct Tile {
    float *W2, *W3, *W1, *b1, *b2, *b3, *x, *h1, *d1, *h2, *d2, *t2;
};

__host__ __device__ inline size_t tile_floats(int O) {
    return (size_t)kH * kHP + (size_t)O * kHP + 3 * kH + O + kT + 5 * (size_t)kT * kHP;
}

__device__ inline Tile carve(float* s, int O) {
    Tile t;
    t.W2 = s;
    t.W3 = t.W2 + kH * kHP;
    t.W1 = t.W3 + O * kHP;
    t.b1 = t.W1 + kH;
    t.b2 = t.b1 + kH;
    t.b3 = t.b2 + kH;
    t.x = t.b3 + O;
    t.h1 = t.x + kT;
    t.d1 = t.h1 + kT * kHP;
    t.h2 = t.d1 + kT * kHP;
    t.d2 = t.h2 + kT * kHP;
    t.t2 = t.d2 + kT * kHP;
    return t;
}

__device__ inline void load_params(const float* p, int O, Tile& t) {
    const MlpOff o = mlp_off(O);
    for (int i = threadIdx.x; i < kH * kH; i += kThr) t.W2[(i / kH) * kHP + (i % kH)] = p[o.W2 + i];
    for (int i = threadIdx.x; i < O * kH; i += kThr) t.W3[(i / kH) * kHP + (i % kH)] = p[o.W3 + i];
    for (int i = threadIdx.x; i < kH; i += kThr) {
        t.W1[i] = p[o.W1 + i];
        t.b1[i] = p[o.b1 + i];
        t.b2[i] = p[o.b2 + i];
    }
    for (int i = threadIdx.x; i < O; i += kThr) t.b3[i] = p[o.b3 + i];
}

// hidden layers of the tile (x already in t.x): h1, d1 = dh1/dx, h2, d2 = dh2/dx, t2 = W2 d1
__device__ inline void forward_hidden(Tile& t) {
    const int tid = threadIdx.x;
    for (int it = tid; it < kT * kH; it += kThr) {
        const int s = it / kH, j = it - s * kH;
        const float h = tanhf(fmaf(t.W1[j], t.x[s], t.b1[j]));
        t.h1[s * kHP + j] = h;
        t.d1[s * kHP + j] = (1.0f - h * h) * t.W1[j];
    }
    __syncthreads();
    const int j = tid & (kH - 1), sg = tid >> 6;
    float a[kT / 4], d[kT / 4];
#pragma unroll
    for (int k = 0; k < kT / 4; ++k) a[k] = d[k] = 0.0f;
    for (int i = 0; i < kH; ++i) {
        const float w = t.W2[j * kHP + i];
#pragma unroll
        for (int k = 0; k < kT / 4; ++k) {
            const int s = sg + 4 * k;
            a[k] = fmaf(w, t.h1[s * kHP + i], a[k]);
            d[k] = fmaf(w, t.d1[s * kHP + i], d[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < kT / 4; ++k) {
        const int s = sg + 4 * k;
        const float h = tanhf(a[k] + t.b2[j]);
        t.h2[s * kHP + j] = h;
        t.d2[s * kHP + j] = (1.0f - h * h) * d[k];
        t.t2[s * kHP + j] = d[k];
    }
    __syncthreads();
}

__device__ inline void output_at(const Tile& t, int s, int o, float& y, float& dy) {
    float a = t.b3[o], d = 0.0f;
    const float* w = t.W3 + o * kHP;
    const float* h = t.h2 + s * kHP;
    const float* g = t.d2 + s * kHP;
    for (int i = 0; i < kH; ++i) {
        a = fmaf(w[i], h[i], a);
        d = fmaf(w[i], g[i], d);
    }
    y = a;
    dy = d;
}

__device__ inline uint32_t mix32(uint32_t h) {
    // murmur3 finaliser
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ inline int sample_index(uint32_t seed, int64_t step, int b, int n_samples) {
    uint32_t h = mix32(seed ^ mix32((uint32_t)step * 0x9e3779b9u + (uint32_t)(step >> 32)));
    h = mix32(h ^ ((uint32_t)b * 0x27d4eb2fu + 0x165667b1u));
    return (int)(h % (uint32_t)n_samples);
}

// ---------------------------------------------------------------------------
__global__ void radial_exact_kernel(cusoap_basis_t b, const double* __restrict__ r, int n_pts,
                                    double* __restrict__ B, double* __restrict__ dB) {
    const int L1 = b.l_max + 1, n = b.n_max;
    const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= (int64_t)n_pts * L1) return;
    const int p = (int)(it / L1), l = (int)(it - (int64_t)p * L1);
    const double rr = r[p], r2 = rr * rr;
    double q[kMaxN], qh[kMaxN];
    for (int k = 0; k < n; ++k) {
        const double g = b.gto_gamma[l * n + k];
        q[k] = b.gto_pref[l * n + k] * exp(-g * r2);
        qh[k] = -2.0 * g * q[k];
    }
    for (int nn = 0; nn < n; ++nn) {
        const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
        double v = 0.0, vh = 0.0;
        for (int k = 0; k < n; ++k) {
            v += beta[k] * q[k];
            vh += beta[k] * qh[k];
        }
        B[((size_t)p * L1 + l) * n + nn] = v;
        dB[((size_t)p * L1 + l) * n + nn] = rr * vh;
    }
}

// y, dy: (n_pairs, n_pts, O); grid (tiles, n_pairs)
__global__ __launch_bounds__(kThr) void mlp_eval_kernel(const float* __restrict__ params, int O,
                                                         const float* __restrict__ x, int n_pts,
                                                         float* __restrict__ y, float* __restrict__ dy) {
    extern __shared__ __align__(16) float sm[];
    Tile t = carve(sm, O);
    const int pair = blockIdx.y, base = blockIdx.x * kT;
    const MlpOff off = mlp_off(O);
    load_params(params + (size_t)pair * off.P, O, t);
    for (int s = threadIdx.x; s < kT; s += kThr) t.x[s] = (base + s < n_pts) ? x[base + s] : 0.0f;
    __syncthreads();
    forward_hidden(t);
    for (int it = threadIdx.x; it < kT * O; it += kThr) {
        const int s = it / O, o = it - s * O;
        if (base + s >= n_pts) continue;
        float v, d;
        output_at(t, s, o, v, d);
        const size_t k = ((size_t)pair * n_pts + base + s) * O + o;
        y[k] = v;
        dy[k] = d;
    }
}

// grid (tiles = batch / kT, n_pairs); partial (n_pairs, tiles, P), ploss (n_pairs, tiles)
__global__ __launch_bounds__(kThr) void mlp_grad_kernel(const float* __restrict__ params, int O,
                                                         const float* __restrict__ xs, const float* __restrict__ ts,
                                                         const float* __restrict__ dts, int n_samples,
                                                         int64_t tstride, const float* __restrict__ wv,
                                                         const float* __restrict__ wd, int batch, uint32_t seed,
                                                         int64_t step, float* __restrict__ partial,
                                                         float* __restrict__ ploss) {
    extern __shared__ __align__(16) float sm[];
    Tile t = carve(sm, O);
    const int OP = O + 1;
    float* ey = t.t2 + kT * kHP;          // [kT][OP]
    float* ed = ey + kT * OP;             // [kT][OP]
    __shared__ int sidx[kT];
    __shared__ float red[kThr / 64];
    const int pair = blockIdx.y, tile = blockIdx.x, tid = threadIdx.x, base = tile * kT;
    const MlpOff off = mlp_off(O);
    load_params(params + (size_t)pair * off.P, O, t);
    if (tid < kT) {
        const int bb = base + tid;
        const int si = bb < batch ? sample_index(seed, step, bb, n_samples) : -1;
        sidx[tid] = si;
        t.x[tid] = si >= 0 ? xs[si] : 0.0f;
    }
    __syncthreads();
    forward_hidden(t);
    // residuals
    const float sc = 1.0f / ((float)O * (float)batch);
    float lacc = 0.0f;
    const float* tp = ts + (size_t)pair * tstride;
    const float* dtp = dts + (size_t)pair * tstride;
    for (int it = tid; it < kT * O; it += kThr) {
        const int s = it / O, o = it - s * O;
        const int si = sidx[s];
        float ry = 0.0f, rd = 0.0f;
        if (si >= 0) {
            float v, d;
            output_at(t, s, o, v, d);
            ry = v - tp[(size_t)si * O + o];
            rd = d - dtp[(size_t)si * O + o];
        }
        lacc += sc * (wv[o] * ry * ry + wd[o] * rd * rd);
        ey[s * OP + o] = 2.0f * sc * wv[o] * ry;
        ed[s * OP + o] = 2.0f * sc * wd[o] * rd;
    }
#pragma unroll
    for (int o2 = 32; o2 > 0; o2 >>= 1) lacc += __shfl_xor(lacc, o2, 64);
    if ((tid & 63) == 0) red[tid >> 6] = lacc;
    __syncthreads();
    float* pg = partial + ((size_t)pair * gridDim.x + tile) * off.P;
    if (tid == 0) ploss[(size_t)pair * gridDim.x + tile] = (red[0] + red[1]) + (red[2] + red[3]);
    // output layer
    for (int it = tid; it < O * kH; it += kThr) {
        const int o = it / kH, i = it - o * kH;
        float g = 0.0f;
        for (int s = 0; s < kT; ++s) g += ey[s * OP + o] * t.h2[s * kHP + i] + ed[s * OP + o] * t.d2[s * kHP + i];
        pg[off.W3 + it] = g;
    }
    for (int o = tid; o < O; o += kThr) {
        float g = 0.0f;
        for (int s = 0; s < kT; ++s) g += ey[s * OP + o];
        pg[off.b3 + o] = g;
    }
    __syncthreads();
    // back through layer 2: ga2 -> h2 slot, gt2 -> t2 slot (each (s, i) touched by one thread)
    {
        const int i = tid & (kH - 1), sg = tid >> 6;
        for (int k = 0; k < kT / 4; ++k) {
            const int s = sg + 4 * k;
            float gh = 0.0f, gd = 0.0f;
            for (int o = 0; o < O; ++o) {
                const float w = t.W3[o * kHP + i];
                gh = fmaf(w, ey[s * OP + o], gh);
                gd = fmaf(w, ed[s * OP + o], gd);
            }
            const float h = t.h2[s * kHP + i], tt = t.t2[s * kHP + i], u = 1.0f - h * h;
            t.t2[s * kHP + i] = u * gd;
            t.h2[s * kHP + i] = u * (gh - 2.0f * h * tt * gd);
        }
    }
    __syncthreads();
    for (int it = tid; it < kH * kH; it += kThr) {
        const int j = it / kH, i = it - j * kH;
        float g = 0.0f;
        for (int s = 0; s < kT; ++s) g += t.h2[s * kHP + j] * t.h1[s * kHP + i] + t.t2[s * kHP + j] * t.d1[s * kHP + i];
        pg[off.W2 + it] = g;
    }
    if (tid < kH) {
        float g = 0.0f;
        for (int s = 0; s < kT; ++s) g += t.h2[s * kHP + tid];
        pg[off.b2 + tid] = g;
    }
    __syncthreads();
    // back through layer 1: ga1 -> h1 slot, (u1 gd1 + ga1 x) -> d1 slot
    {
        const int i = tid & (kH - 1), sg = tid >> 6;
        const float w1 = t.W1[i];
        for (int k = 0; k < kT / 4; ++k) {
            const int s = sg + 4 * k;
            float gh = 0.0f, gd = 0.0f;
            for (int j = 0; j < kH; ++j) {
                const float w = t.W2[j * kHP + i];
                gh = fmaf(w, t.h2[s * kHP + j], gh);
                gd = fmaf(w, t.t2[s * kHP + j], gd);
            }
            const float h = t.h1[s * kHP + i], u = 1.0f - h * h;
            const float ga1 = u * (gh - 2.0f * h * w1 * gd);
            t.h1[s * kHP + i] = ga1;
            t.d1[s * kHP + i] = u * gd + ga1 * t.x[s];
        }
    }
    __syncthreads();
    if (tid < kH) {
        float gw = 0.0f, gb = 0.0f;
        for (int s = 0; s < kT; ++s) {
            gw += t.d1[s * kHP + tid];
            gb += t.h1[s * kHP + tid];
        }
        pg[off.W1 + tid] = gw;
        pg[off.b1 + tid] = gb;
    }
}

// grad (n_pairs, P) = fixed-order sum over tiles; loss (n_pairs)
__global__ void mlp_reduce_kernel(const float* __restrict__ partial, const float* __restrict__ ploss, int tiles,
                                  int P, int n_pairs, float* __restrict__ grad, float* __restrict__ loss) {
    const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it < (int64_t)n_pairs * P) {
        const int pair = (int)(it / P), p = (int)(it - (int64_t)pair * P);
        const float* src = partial + (size_t)pair * tiles * P + p;
        float g = 0.0f;
        for (int k = 0; k < tiles; ++k) g += src[(size_t)k * P];
        grad[it] = g;
    }
    if (loss && it < n_pairs) {
        float l = 0.0f;
        for (int k = 0; k < tiles; ++k) l += ploss[(size_t)it * tiles + k];
        loss[it] = l;
    }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float beta1, float beta2, float eps, float c1,
                            float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// rows (n_pairs, 2, n_samples, RL), RL = kH + 1 + O:  [h2, 1, t] and [dh2/dx, 0, dt]
__global__ __launch_bounds__(kThr) void mlp_rows_kernel(const float* __restrict__ params, int O,
                                                         const float* __restrict__ xs, const float* __restrict__ ts,
                                                         const float* __restrict__ dts, int n_samples,
                                                         int64_t tstride, float* __restrict__ rows) {
    extern __shared__ __align__(16) float sm[];
    Tile t = carve(sm, O);
    const int pair = blockIdx.y, base = blockIdx.x * kT, RL = kH + 1 + O;
    const MlpOff off = mlp_off(O);
    load_params(params + (size_t)pair * off.P, O, t);
    for (int s = threadIdx.x; s < kT; s += kThr) t.x[s] = (base + s < n_samples) ? xs[base + s] : 0.0f;
    __syncthreads();
    forward_hidden(t);
    float* U = rows + (size_t)pair * 2 * n_samples * RL;
    float* V = U + (size_t)n_samples * RL;
    const float* tp = ts + (size_t)pair * tstride;
    const float* dtp = dts + (size_t)pair * tstride;
    for (int it = threadIdx.x; it < kT * RL; it += kThr) {
        const int s = it / RL, c = it - s * RL;
        const int g = base + s;
        if (g >= n_samples) continue;
        float u, v;
        if (c < kH) {
            u = t.h2[s * kHP + c];
            v = t.d2[s * kHP + c];
        } else if (c == kH) {
            u = 1.0f;
            v = 0.0f;
        } else {
            u = tp[(size_t)g * O + (c - kH - 1)];
            v = dtp[(size_t)g * O + (c - kH - 1)];
        }
        U[(size_t)g * RL + c] = u;
        V[(size_t)g * RL + c] = v;
    }
}

// normal-equation entries: side (U, V) x row a < kH+1 x column c < RL, summed over one chunk
// grid (entry blocks, n_pairs, chunks); part (n_pairs, chunks, n_entries) float64
constexpr int kEPT = 4;
__global__ __launch_bounds__(kThr) void mlp_normal_kernel(const float* __restrict__ rows, int O, int n_samples,
                                                           int chunk, double* __restrict__ part) {
    extern __shared__ __align__(16) float sm[];
    const int RL = kH + 1 + O, n_ent = 2 * kHP * RL;
    float* Us = sm;                     // [kT][RL]
    float* Vs = Us + kT * RL;           // [kT][RL]
    const int pair = blockIdx.y, ch = blockIdx.z, tid = threadIdx.x;
    const float* U = rows + (size_t)pair * 2 * n_samples * RL;
    const float* V = U + (size_t)n_samples * RL;
    int side[kEPT], ra[kEPT], cb[kEPT];
    double acc[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const int e = (blockIdx.x * kEPT + q) * kThr + tid;
        acc[q] = 0.0;
        if (e < n_ent) {
            side[q] = e / (kHP * RL);
            const int rem = e - side[q] * kHP * RL;
            ra[q] = rem / RL;
            cb[q] = rem - ra[q] * RL;
        } else {
            side[q] = -1;
            ra[q] = cb[q] = 0;
        }
    }
    const int s0 = ch * chunk, s1 = min(n_samples, s0 + chunk);
    for (int sb = s0; sb < s1; sb += kT) {
        const int ns = min(kT, s1 - sb);
        for (int i = tid; i < ns * RL; i += kThr) {
            Us[i] = U[(size_t)sb * RL + i];
            Vs[i] = V[(size_t)sb * RL + i];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kEPT; ++q) {
            if (side[q] < 0) continue;
            const float* X = side[q] ? Vs : Us;
            double a = 0.0;
            for (int s = 0; s < ns; ++s) a += (double)X[s * RL + ra[q]] * (double)X[s * RL + cb[q]];
            acc[q] += a;
        }
        __syncthreads();
    }
    double* dst = part + ((size_t)pair * gridDim.z + ch) * n_ent;
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const int e = (blockIdx.x * kEPT + q) * kThr + tid;
        if (e < n_ent) dst[e] = acc[q];
    }
}

// one workgroup per (output channel, pair): A = AU + wd_o AV + ridge, rhs = RU + wd_o RV -> W3[o], b3[o]
__global__ __launch_bounds__(kThr) void mlp_solve_kernel(const double* __restrict__ part, int O, int chunks,
                                                          const float* __restrict__ wd, double ridge,
                                                          float* __restrict__ params) {
    __shared__ double A[kHP * kHP];
    __shared__ double rhs[kHP];
    __shared__ double diag_sum;
    const int o = blockIdx.x, pair = blockIdx.y, tid = threadIdx.x;
    const int RL = kH + 1 + O, n_ent = 2 * kHP * RL;
    const double w = (double)wd[o];
    const double* base = part + (size_t)pair * chunks * n_ent;
    for (int it = tid; it < kHP * (kHP + 1); it += kThr) {
        const int a = it / (kHP + 1), c = it - a * (kHP + 1);
        const int col = c < kHP ? c : kHP + o;        // last column: the target of channel o
        double u = 0.0, v = 0.0;
        for (int k = 0; k < chunks; ++k) {
            u += base[(size_t)k * n_ent + a * RL + col];
            v += base[(size_t)k * n_ent + kHP * RL + a * RL + col];
        }
        if (c < kHP) A[a * kHP + c] = u + w * v;
        else rhs[a] = u + w * v;
    }
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int a = 0; a < kHP; ++a) tr += A[a * kHP + a];
        diag_sum = tr;
    }
    __syncthreads();
    if (tid < kHP) A[tid * kHP + tid] += ridge * diag_sum / kHP;
    __syncthreads();
    // right-looking Cholesky, lower triangle in place
    for (int k = 0; k < kHP; ++k) {
        if (tid == 0) A[k * kHP + k] = sqrt(fmax(A[k * kHP + k], 1e-300));
        __syncthreads();
        const double dkk = A[k * kHP + k];
        for (int i = k + 1 + tid; i < kHP; i += kThr) A[i * kHP + k] /= dkk;
        __syncthreads();
        for (int it = tid; it < (kHP - k - 1) * (kHP - k - 1); it += kThr) {
            const int i = k + 1 + it / (kHP - k - 1), j = k + 1 + it % (kHP - k - 1);
            if (j <= i) A[i * kHP + j] -= A[i * kHP + k] * A[j * kHP + k];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int i = 0; i < kHP; ++i) {           // L z = rhs
            double v = rhs[i];
            for (int j = 0; j < i; ++j) v -= A[i * kHP + j] * rhs[j];
            rhs[i] = v / A[i * kHP + i];
        }
        for (int i = kHP - 1; i >= 0; --i) {      // L^T w = z
            double v = rhs[i];
            for (int j = i + 1; j < kHP; ++j) v -= A[j * kHP + i] * rhs[j];
            rhs[i] = v / A[i * kHP + i];
        }
    }
    __syncthreads();
    const MlpOff off = mlp_off(O);
    float* p = params + (size_t)pair * off.P;
    if (tid < kH) p[off.W3 + o * kH + tid] = (float)rhs[tid];
    if (tid == 0) p[off.b3 + o] = (float)rhs[kH];
}

__global__ void node_inputs_kernel(int n_nodes, double h, double r_fit, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const double r = fmin(i * h, r_fit);
    x[i] = (float)(2.0 * r / r_fit - 1.0);
}

// table ((S+1) S slots, G+1 nodes, L1, n, 2): (B, dB/dr) pairs, so a kernel reads
// one channel's node with one 8-byte load; slot = center_slot * S + neighbour
// species, center_slot = S for centers that are not atoms (always the exact function)
__global__ void table_kernel(cusoap_basis_t b, const float* __restrict__ y, const float* __restrict__ dy,
                             const int32_t* __restrict__ use_pair, const float* __restrict__ scale, double r_fit,
                             int G, double h, float* __restrict__ table) {
    const int S = b.n_species, L1 = b.l_max + 1, n = b.n_max, O = n * L1;
    const int n_nodes = G + 1;
    const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)(S + 1) * S * n_nodes * L1;
    if (it >= total) return;
    const int l = (int)(it % L1);
    const int node = (int)((it / L1) % n_nodes);
    const int slot = (int)(it / ((int64_t)L1 * n_nodes));
    const int cs = slot / S;
    const double r = node * h;
    const bool net = cs < S && use_pair && use_pair[slot] && node > 0 && r <= r_fit * (1.0 + 1e-12);
    float* dst = table + (((size_t)slot * n_nodes + node) * L1 + l) * 2 * n;
    if (net) {
        const int pair = slot;               // slots below S*S are the species pairs
        const double u = r / r_fit;
        double ul = 1.0;
        for (int k = 0; k < l; ++k) ul *= u;
        for (int nn = 0; nn < n; ++nn) {
            const int j = l * n + nn;
            const size_t k = ((size_t)pair * n_nodes + node) * O + j;
            const double sj = (double)scale[(size_t)pair * O + j];
            const double g = sj * (double)y[k], gp = sj * (2.0 / r_fit) * (double)dy[k];
            dst[2 * nn] = (float)(g / ul);
            dst[2 * nn + 1] = (float)(gp / ul - l * g / (r_fit * ul * u));
        }
        return;
    }
    const double r2 = r * r;
    double q[kMaxN];
    for (int k = 0; k < n; ++k) q[k] = b.gto_pref[l * n + k] * exp(-b.gto_gamma[l * n + k] * r2);
    for (int nn = 0; nn < n; ++nn) {
        const double* beta = b.gto_beta + ((size_t)l * n + nn) * n;
        double v = 0.0, vh = 0.0;
        for (int k = 0; k < n; ++k) {
            v += beta[k] * q[k];
            vh += beta[k] * (-2.0 * b.gto_gamma[l * n + k]) * q[k];
        }
        dst[2 * nn] = (float)v;
        dst[2 * nn + 1] = (float)(r * vh);
    }
}

int check_net(int n_hidden, int O) {
    if (n_hidden != kH) return fail(CUSOAP_ERR_UNSUPPORTED, "mlp: hidden width must be 64");
    if (O < 1 || O > kMaxO) return fail(CUSOAP_ERR_UNSUPPORTED, "mlp: output count must be 1..128");
    return CUSOAP_OK;
}

template <typename K>
int set_smem(K kern, size_t bytes) {
    if (bytes > 160 * 1024) return fail(CUSOAP_ERR_UNSUPPORTED, "mlp: tile does not fit in LDS");
    if (bytes > 64 * 1024)
        CUSOAP_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return CUSOAP_OK;
}

}  // namespace

int64_t mlp_param_count(int n_hidden, int n_out) {
    if (n_hidden != kH || n_out < 1 || n_out > kMaxO) return -1;
    return mlp_off(n_out).P;
}

int radial_exact_impl(const cusoap_basis_t* b, const double* r, int n_pts, double* B, double* dB, hipStream_t st) {
    if (!b || b->rbf != CUSOAP_RBF_GTO) return fail(CUSOAP_ERR_UNSUPPORTED, "radial_exact: GTO basis only");
    if (b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "radial_exact: n_max must be 1..16 and l_max 0..9");
    if (n_pts < 0) return fail(CUSOAP_ERR_ARG, "radial_exact: negative size");
    if (n_pts == 0) return CUSOAP_OK;
    const int64_t items = (int64_t)n_pts * (b->l_max + 1);
    radial_exact_kernel<<<(unsigned)((items + 255) / 256), 256, 0, st>>>(*b, r, n_pts, B, dB);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

int mlp_eval_impl(const float* params, int n_pairs, int n_hidden, int O, const float* x, int n_pts, float* y,
                  float* dy, hipStream_t st) {
    if (int rc = check_net(n_hidden, O)) return rc;
    if (n_pairs < 1 || n_pts < 0) return fail(CUSOAP_ERR_ARG, "mlp_eval: bad sizes");
    if (n_pts == 0) return CUSOAP_OK;
    const size_t smem = sizeof(float) * tile_floats(O);
    if (int rc = set_smem(mlp_eval_kernel, smem)) return rc;
    mlp_eval_kernel<<<dim3((n_pts + kT - 1) / kT, n_pairs), kThr, smem, st>>>(params, O, x, n_pts, y, dy);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

size_t mlp_train_ws_bytes(int n_pairs, int O, int batch) {
    const int tiles = (batch + kT - 1) / kT;
    return align_up(sizeof(float) * (size_t)n_pairs * tiles * (mlp_off(O).P + 1), 256);
}

int mlp_loss_grad_impl(const float* params, int n_pairs, int n_hidden, int O, const float* x, const float* t,
                       const float* dt, int n_samples, int64_t tstride, const float* wv, const float* wd, int batch,
                       uint32_t seed, int64_t step, void* ws, size_t ws_bytes, float* grad, float* loss,
                       hipStream_t st) {
    if (int rc = check_net(n_hidden, O)) return rc;
    if (n_pairs < 1 || n_samples < 1 || batch < 1) return fail(CUSOAP_ERR_ARG, "mlp_loss_grad: bad sizes");
    if (ws_bytes < mlp_train_ws_bytes(n_pairs, O, batch))
        return fail(CUSOAP_ERR_WORKSPACE, "mlp_loss_grad: workspace too small");
    const int tiles = (batch + kT - 1) / kT;
    const int P = mlp_off(O).P;
    float* partial = (float*)ws;
    float* ploss = partial + (size_t)n_pairs * tiles * P;
    const size_t smem = sizeof(float) * (tile_floats(O) + 2 * (size_t)kT * (O + 1));
    if (int rc = set_smem(mlp_grad_kernel, smem)) return rc;
    mlp_grad_kernel<<<dim3(tiles, n_pairs), kThr, smem, st>>>(params, O, x, t, dt, n_samples, tstride, wv, wd, batch,
                                                              seed, step, partial, ploss);
    CUSOAP_LAUNCH_CHECK();
    const int64_t items = (int64_t)n_pairs * P;
    mlp_reduce_kernel<<<(unsigned)((items + 255) / 256), 256, 0, st>>>(partial, ploss, tiles, P, n_pairs, grad, loss);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

int adam_impl(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
              float eps, int64_t step, hipStream_t st) {
    if (n < 0 || step < 1) return fail(CUSOAP_ERR_ARG, "adam: n >= 0 and step >= 1 required");
    if (n == 0) return CUSOAP_OK;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)step)));
    adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

namespace {
constexpr int kRefitChunk = 2048;
}

size_t mlp_refit_ws_bytes(int n_pairs, int O, int n_samples) {
    const int RL = kH + 1 + O, n_ent = 2 * kHP * RL;
    const int chunks = (n_samples + kRefitChunk - 1) / kRefitChunk;
    return align_up(sizeof(float) * (size_t)n_pairs * 2 * n_samples * RL, 256) +
           align_up(sizeof(double) * (size_t)n_pairs * chunks * n_ent, 256);
}

int mlp_refit_impl(float* params, int n_pairs, int n_hidden, int O, const float* x, const float* t, const float* dt,
                   int n_samples, int64_t tstride, const float* wd, double ridge, void* ws, size_t ws_bytes,
                   hipStream_t st) {
    if (int rc = check_net(n_hidden, O)) return rc;
    if (n_pairs < 1 || n_samples < kHP) return fail(CUSOAP_ERR_ARG, "mlp_refit: need at least 65 samples");
    if (ws_bytes < mlp_refit_ws_bytes(n_pairs, O, n_samples))
        return fail(CUSOAP_ERR_WORKSPACE, "mlp_refit: workspace too small");
    const int RL = kH + 1 + O, n_ent = 2 * kHP * RL;
    const int chunks = (n_samples + kRefitChunk - 1) / kRefitChunk;
    float* rows = (float*)ws;
    double* part = (double*)((char*)ws + align_up(sizeof(float) * (size_t)n_pairs * 2 * n_samples * RL, 256));
    const size_t smem_rows = sizeof(float) * tile_floats(O);
    if (int rc = set_smem(mlp_rows_kernel, smem_rows)) return rc;
    mlp_rows_kernel<<<dim3((n_samples + kT - 1) / kT, n_pairs), kThr, smem_rows, st>>>(params, O, x, t, dt, n_samples,
                                                                                       tstride, rows);
    CUSOAP_LAUNCH_CHECK();
    const size_t smem_n = sizeof(float) * 2 * (size_t)kT * RL;
    if (int rc = set_smem(mlp_normal_kernel, smem_n)) return rc;
    const int eblocks = (n_ent + kThr * kEPT - 1) / (kThr * kEPT);
    mlp_normal_kernel<<<dim3(eblocks, n_pairs, chunks), kThr, smem_n, st>>>(rows, O, n_samples, kRefitChunk, part);
    CUSOAP_LAUNCH_CHECK();
    mlp_solve_kernel<<<dim3(O, n_pairs), kThr, 0, st>>>(part, O, chunks, wd, ridge, params);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

size_t radial_table_ws_bytes(const cusoap_basis_t* b, int n_intervals) {
    const int S = b->n_species, O = b->n_max * (b->l_max + 1), nodes = n_intervals + 1;
    return align_up(sizeof(float) * (size_t)nodes, 256) + 2 * align_up(sizeof(float) * (size_t)S * S * nodes * O, 256);
}

int radial_table_impl(const cusoap_basis_t* b, const float* params, int n_hidden, const int32_t* use_pair,
                      const float* scale, double r_fit, int n_intervals, double r_max, void* ws, size_t ws_bytes,
                      float* table, hipStream_t st) {
    if (!b || b->rbf != CUSOAP_RBF_GTO) return fail(CUSOAP_ERR_UNSUPPORTED, "radial_table: GTO basis only");
    if (b->n_max < 1 || b->n_max > kMaxN || b->l_max < 0 || b->l_max > kMaxL)
        return fail(CUSOAP_ERR_UNSUPPORTED, "radial_table: n_max must be 1..16 and l_max 0..9");
    if (n_intervals < 1 || !(r_max > 0.0) || !(r_fit > 0.0)) return fail(CUSOAP_ERR_ARG, "radial_table: bad grid");
    const int S = b->n_species, O = b->n_max * (b->l_max + 1), nodes = n_intervals + 1;
    const double h = r_max / n_intervals;
    float *y = nullptr, *dy = nullptr;
    if (params) {
        if (int rc = check_net(n_hidden, O)) return rc;
        if (!use_pair || !scale) return fail(CUSOAP_ERR_ARG, "radial_table: use_pair and scale required");
        if (ws_bytes < radial_table_ws_bytes(b, n_intervals))
            return fail(CUSOAP_ERR_WORKSPACE, "radial_table: workspace too small");
        float* x = (float*)ws;
        y = (float*)((char*)ws + align_up(sizeof(float) * (size_t)nodes, 256));
        dy = (float*)((char*)y + align_up(sizeof(float) * (size_t)S * S * nodes * O, 256));
        node_inputs_kernel<<<(nodes + 255) / 256, 256, 0, st>>>(nodes, h, r_fit, x);
        CUSOAP_LAUNCH_CHECK();
        if (int rc = mlp_eval_impl(params, S * S, n_hidden, O, x, nodes, y, dy, st)) return rc;
    }
    const int64_t items = (int64_t)(S + 1) * S * nodes * (b->l_max + 1);
    table_kernel<<<(unsigned)((items + 255) / 256), 256, 0, st>>>(*b, y, dy, params ? use_pair : nullptr, scale, r_fit,
                                                                  n_intervals, h, table);
    CUSOAP_LAUNCH_CHECK();
    return CUSOAP_OK;
}

}  // namespace cusoap
